/*
 * phylo_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * FP64 CPU restatement of phylommand treeator's Felsenstein-pruning path.
 * It is the checker for the HIP library: only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it, and never as the thing that is
 * measured or shipped.  Every function cites the reference file:line it
 * restates (paths relative to the phylommand source tree).
 *
 * Pinning (what checks each function against something other than itself):
 *   orc_tree_parse          golden newick fixtures written by the compiled
 *                           reference (tree.cpp tree_file_parser + print_newick)
 *   orc_parse_fasta         golden per-site lnL through the reference's own
 *                           matrix_parser (same fasta in, same lnL out)
 *   orc_build_Q             bit-exact golden lnL (refcompat P) for <=3-rate
 *                           models; closed-form JC69/K80 P(t)
 *   orc_pmatrix REFCOMPAT   bit-exact golden P(t) dumps from sub_model::set_P_matrix
 *   orc_pmatrix EIG/PADE    scipy.linalg.expm, P(s)P(t)=P(s+t), rows sum to 1,
 *                           JC69 closed form
 *   orc_time_mod_length     golden get_time_mod_branch_length dumps
 *   orc_effective_lengths   golden lnL with time cut-offs / clade rates
 *   orc_gamma_rates         scipy.special.gammaincinv/gammainc, Yang (1994)
 *                           table values, mean rate == 1
 *   orc_loglik              golden per-site lnL (bit-exact, refcompat, C=1),
 *                           brute-force state enumeration on tiny trees,
 *                           two-taxon closed form, pulley principle
 *   orc_node_partials       golden [&P={...}] labels from
 *                           draw_normalized_likelihood_on_nodes
 *   orc_nni_reference       golden treebender --nni all tree lists
 *   orc_nni_list mode 0     the same golden lists (variant order included)
 *   orc_nni_list mode 1     brute force on split sets: the variants are the
 *                           2(n-3) distinct unrooted trees at Robinson-Foulds
 *                           distance 2, for every topology with n = 5..8
 *   orc_simulate            the reference's draws cannot be replayed (libc
 *                           rand() seeded by time(NULL), tree.h:61-64); its
 *                           distribution is pinned: chi-square of 4-taxon
 *                           pattern counts against exp(site lnL), with
 *                           P = I, swapped lengths, ignored categories and
 *                           wrong root frequencies all rejected
 */
#include <ctype.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define ORC_OK 0
#define ORC_ERR -1

/* ------------------------------------------------------------------ */
/* Tree: newick parser restating tree::tree_file_parser (tree.cpp:69-237) */
/* ------------------------------------------------------------------ */

typedef struct {
    int n, cap;
    int root;
    int *left, *right, *parent;
    double *blen;
    char **label;
} orc_tree;

static int tree_new_node(orc_tree *t) {
    if (t->n == t->cap) {
        int nc = t->cap ? 2 * t->cap : 64;
        t->left = realloc(t->left, nc * sizeof(int));
        t->right = realloc(t->right, nc * sizeof(int));
        t->parent = realloc(t->parent, nc * sizeof(int));
        t->blen = realloc(t->blen, nc * sizeof(double));
        t->label = realloc(t->label, nc * sizeof(char *));
        t->cap = nc;
    }
    int k = t->n++;
    t->left[k] = t->right[k] = t->parent[k] = -1;
    t->blen[k] = 0.0;
    t->label[k] = NULL;
    return k;
}

void orc_tree_free(orc_tree *t) {
    if (!t) return;
    for (int i = 0; i < t->n; ++i) free(t->label[i]);
    free(t->left); free(t->right); free(t->parent); free(t->blen); free(t->label);
    free(t);
}

typedef struct { char *s; size_t n, cap; } sbuf;
static void sb_push(sbuf *b, char c) {
    if (b->n + 2 > b->cap) { b->cap = b->cap ? 2 * b->cap : 32; b->s = realloc(b->s, b->cap); }
    b->s[b->n++] = c; b->s[b->n] = 0;
}
static void sb_clear(sbuf *b) { b->n = 0; if (b->s) b->s[0] = 0; }

/* `infile >> temp` skips whitespace (tree.cpp:81,235) */
static int nextc(const char *s, size_t *p, char *c) {
    while (s[*p] && isspace((unsigned char)s[*p])) (*p)++;
    if (!s[*p]) return 0;
    *c = s[(*p)++];
    return 1;
}

static void assign_label_len(orc_tree *t, int node, sbuf *lab, sbuf *bl) {
    /* tree.cpp:128-140 (','), :164-176 (')'), :182-193 (';') */
    if (lab->n) { free(t->label[node]); t->label[node] = strdup(lab->s); sb_clear(lab); }
    if (bl->n) { t->blen[node] = atof(bl->s); sb_clear(bl); }
}

/* Parses one newick tree from s starting at *pos.  Returns a tree with
 * n == 0 when nothing was parsed or brackets were unbalanced (tree.cpp:198-202). */
orc_tree *orc_tree_parse_at(const char *s, size_t *pos) {
    orc_tree *t = calloc(1, sizeof(orc_tree));
    t->root = tree_new_node(t);
    int present = t->root;
    sbuf lab = {0}, bl = {0};
    char mode = 's', c;
    int ok = nextc(s, pos, &c);
    while (ok) {
        if (c == '[') {                       /* comments, tree.cpp:84-107 */
            ok = nextc(s, pos, &c);
            if (!ok) break;
            if (c != ']') {
                int depth = 1;
                while (ok && depth > 0) {
                    if (c == '[') ++depth;
                    if (c == ']') --depth;
                    ok = nextc(s, pos, &c);
                }
                continue;
            }
        } else if (c == '(') {                /* tree.cpp:109-118 */
            mode = 'l';
            int k = tree_new_node(t);
            t->left[present] = k; t->parent[k] = present; present = k;
            sb_clear(&lab);
        } else if (c == ',') {                /* tree.cpp:120-160 */
            mode = 'l';
            assign_label_len(t, present, &lab, &bl);
            present = t->parent[present];
            if (present < 0) { t->n = 0; break; }
            if (t->right[present] < 0) {
                int k = tree_new_node(t);
                t->right[present] = k; t->parent[k] = present; present = k;
            } else {                          /* polytomy: zero-length node, tree.cpp:146-158 */
                int e = tree_new_node(t);
                t->right[e] = t->right[present];
                t->left[e] = t->left[present];
                t->parent[t->right[e]] = e;
                t->parent[t->left[e]] = e;
                t->parent[e] = present;
                t->left[present] = e;
                int k = tree_new_node(t);
                t->right[present] = k; t->parent[k] = present; present = k;
            }
        } else if (c == ')') {                /* tree.cpp:162-179 */
            mode = 'l';
            assign_label_len(t, present, &lab, &bl);
            present = t->parent[present];
            if (present < 0) { t->n = 0; break; }
        } else if (c == ':') {
            mode = 'b';
        } else if (c == ';') {                /* tree.cpp:181-205 */
            if (present == t->root) {
                assign_label_len(t, present, &lab, &bl);
                int r = t->root;
                if (t->right[r] < 0 && t->left[r] >= 0 && t->label[r]) {   /* rooted on a leaf */
                    int k = tree_new_node(t);
                    t->right[r] = k; t->parent[k] = r;
                    t->label[k] = t->label[r]; t->label[r] = NULL;
                    t->blen[k] = t->blen[r]; t->blen[r] = 0.0;
                }
            } else {
                t->n = 0;                     /* unbalanced: tree destroyed */
            }
            break;
        } else if (mode == 'b') {
            sb_push(&bl, c);
        } else if (mode == 'l') {             /* tree.cpp:212-229, quoted labels */
            if (lab.n == 0 && (c == '\'' || c == '"')) {
                char marker = c;
                ok = nextc(s, pos, &c);
                while (ok) {
                    if (c == marker) break;
                    sb_push(&lab, c);
                    ok = nextc(s, pos, &c);
                }
            } else {
                sb_push(&lab, c);
            }
        }
        ok = nextc(s, pos, &c);
    }
    free(lab.s); free(bl.s);
    return t;
}

orc_tree *orc_tree_parse(const char *s) { size_t p = 0; return orc_tree_parse_at(s, &p); }
int orc_tree_n_nodes(const orc_tree *t) { return t->n; }
int orc_tree_root(const orc_tree *t) { return t->root; }
void orc_tree_arrays(const orc_tree *t, int *left, int *right, int *parent, double *blen) {
    for (int i = 0; i < t->n; ++i) {
        left[i] = t->left[i]; right[i] = t->right[i]; parent[i] = t->parent[i]; blen[i] = t->blen[i];
    }
}
const char *orc_tree_label(const orc_tree *t, int i) { return (i >= 0 && i < t->n && t->label[i]) ? t->label[i] : ""; }

/* print_newick_subtree with labels and 6-decimal fixed lengths (tree.cpp:239-280). */
static void pn(const orc_tree *t, int v, sbuf *o, int with_len) {
    char buf[64];
    if (t->left[v] >= 0) { sb_push(o, '('); pn(t, t->left[v], o, with_len); }
    else if (t->label[v]) for (const char *q = t->label[v]; *q; ++q) sb_push(o, *q);
    if (t->right[v] >= 0) {
        sb_push(o, ','); pn(t, t->right[v], o, with_len); sb_push(o, ')');
        if (t->label[v]) for (const char *q = t->label[v]; *q; ++q) sb_push(o, *q);
    }
    if (v != t->root && with_len) {
        snprintf(buf, sizeof buf, ":%f", t->blen[v]);
        for (const char *q = buf; *q; ++q) sb_push(o, *q);
    }
}
/* Writes newick (";" terminated) into out; returns the needed length. */
int orc_tree_newick(const orc_tree *t, char *out, int cap) {
    sbuf o = {0};
    pn(t, t->root, &o, 1);
    sb_push(&o, ';');
    int n = (int)o.n;
    if (out && cap > n) memcpy(out, o.s, n + 1);
    free(o.s);
    return n;
}

/* ------------------------------------------------------------------ */
/* Alphabets and fasta (matrix_parser.cpp:93-139, :274-383)            */
/* ------------------------------------------------------------------ */

/* alphabet::set_alphabet_dna (matrix_parser.cpp:282-323): A=0 G=1 C=2 T=3 */
static uint32_t dna_mask(char c) {
    switch (toupper((unsigned char)c)) {
    case 'A': return 1u; case 'G': return 2u; case 'C': return 4u; case 'T': return 8u;
    case 'R': return 1u | 2u; case 'Y': return 4u | 8u; case 'S': return 2u | 4u;
    case 'W': return 1u | 8u; case 'K': return 2u | 8u; case 'M': return 1u | 4u;
    case 'B': return 2u | 4u | 8u; case 'D': return 1u | 2u | 8u; case 'H': return 1u | 4u | 8u;
    case 'V': return 1u | 2u | 4u; case 'N': return 15u;
    default: return 0u;                /* '-', '.', unknown: empty (map::operator[], :129) */
    }
}
/* alphabet::set_alphabet_amino_acid (matrix_parser.cpp:324-375): 23 letters */
static uint32_t aa_mask(char c) {
    static const char order[] = "ABCDEFGHIKLMNPQRSTVWXYZ";
    char u = (char)toupper((unsigned char)c);
    const char *p = strchr(order, u);
    if (!p || !u) return 0u;
    return 1u << (p - order);
}
/* alphabet::set_alphabet_binary (matrix_parser.cpp:275-281) */
static uint32_t bin_mask(char c) {
    if (c == '0') return 1u; if (c == '1') return 2u; if (c == '-') return 3u;
    return 0u;
}

/* A whitespace alphabet file (alphabet_parser::pars_whitespace,
 * matrix_parser.cpp:227-272) as a 256-entry character table. */
static uint32_t custom_table[256];
void orc_set_custom_alphabet(const uint32_t *table) { memcpy(custom_table, table, sizeof custom_table); }

/* alphabet: 0 dna, 1 protein, 2 binary, 3 custom table. */
uint32_t orc_alphabet_mask(int alphabet, char c) {
    if (alphabet == 3) return custom_table[(unsigned char)c];
    return alphabet == 1 ? aa_mask(c) : alphabet == 2 ? bin_mask(c) : dna_mask(c);
}
/* Union of every letter's states (treeator.cpp:424-426). */
uint32_t orc_alphabet_all(int alphabet) {
    if (alphabet == 3) { uint32_t a = 0; for (int i = 0; i < 256; ++i) a |= custom_table[i]; return a; }
    return alphabet == 1 ? ((1u << 23) - 1u) : alphabet == 2 ? 3u : 15u;
}

/* Parses a fasta text following matrix_parser::pars_fasta (matrix_parser.cpp:93-139),
 * including the extra read of the final character when the text does not end
 * in a newline (the `while (file) { file.get(c); ...}` loop body runs once
 * more with the stale character), then set_empty_to(all states)
 * (character_vector.h:83-86 via treeator.cpp:463-466).
 * Output: names (NUL-separated into names_out), masks [taxon][max_sites].
 * Returns number of taxa; *n_sites_out = max row length. Call with NULL
 * outputs to size the buffers. */
int orc_parse_fasta(const char *text, int alphabet, int set_empty,
                    char *names_out, int names_cap, uint32_t *masks_out, int max_sites,
                    int *row_len_out, int *n_sites_out) {
    int n_taxa = 0, pos = 0, max_len = 0, row_open = 0, row_has = 0, ntaxa_rows = 0;
    int names_used = 0;
    char mode = 'n';
    sbuf taxon = {0};
    size_t len = strlen(text);
    uint32_t all = orc_alphabet_all(alphabet);
    int cur_row = -1;
    for (size_t i = 0; i <= len; ++i) {
        char c;
        if (i < len) c = text[i];
        else { if (len == 0) break; c = text[len - 1]; }   /* stale re-read at EOF */
        if (c == '>') mode = 't';
        else if (mode == 't' && (c == '\n' || c == '\r')) {
            ++n_taxa;
            cur_row = ntaxa_rows++;
            row_open = 1; row_has = 0;
            if (names_out) {
                int need = (int)taxon.n + 1;
                if (names_used + need <= names_cap) {
                    memcpy(names_out + names_used, taxon.n ? taxon.s : "", need);
                }
                names_used += need;
            } else names_used += (int)taxon.n + 1;
            sb_clear(&taxon);
            pos = 0;
            mode = 's';
            if (row_len_out) row_len_out[cur_row] = 0;
        } else if (c != '\n' && c != '\r' && c != '\t' && c != ' ') {
            if (mode == 't') sb_push(&taxon, c);
            else if (mode == 's' && row_open) {
                uint32_t m = orc_alphabet_mask(alphabet, c);
                if (set_empty && m == 0) m = all;
                if (masks_out && pos < max_sites) masks_out[(size_t)cur_row * max_sites + pos] = m;
                ++pos; row_has = 1;
                if (row_len_out) row_len_out[cur_row] = pos;
                if (pos > max_len) max_len = pos;
            }
        }
    }
    (void)row_has;
    free(taxon.s);
    if (n_sites_out) *n_sites_out = max_len;
    return n_taxa;
}

/* n_states = highest set state + 1 over all taxa (treeator.cpp:555-558,
 * character_vector.h:75-82). */
int orc_n_states(const uint32_t *masks, long n) {
    int hi = 0;
    for (long i = 0; i < n; ++i)
        for (int j = hi; j < 32; ++j) if (masks[i] & (1u << j)) hi = j;
    return hi + 1;
}

/* ------------------------------------------------------------------ */
/* sub_model: Q construction (sub_model.cpp:3-53, :99-114)             */
/* ------------------------------------------------------------------ */

/* get_rate_pos (sub_model.cpp:3-29) */
int orc_rate_pos(int S, int tr, int row, int col) {
    int npos = tr ? (S * S - S) / 2 : S * S - S;
    if (row == col) return npos;
    int pos = 0;
    if (tr) {
        int i = 0, j = 0;
        if (col > row) { i = row; j = col - row; }
        else { i = col; j = row - col; }
        while (i) { pos += S - i; --i; }
        pos += j - 1;
    } else {
        for (int i = 0; i < row; ++i) pos += S - 1;
        pos += col;
        if (col > row) --pos;
    }
    return pos;
}

/* get_freq (sub_model.cpp:44-53): the last frequency is 1 - sum of the others,
 * accumulated in index order.  freqs may be NULL (equal frequencies). */
void orc_freqs(int S, const double *freqs_in, double *out) {
    for (int n = 0; n < S; ++n) {
        if (!freqs_in) { out[n] = 1.0 / S; continue; }
        if (n == S - 1) {
            double f = 1.0;
            for (int i = 0; i < S - 1; ++i) f -= freqs_in[i];
            out[n] = f;
        } else out[n] = freqs_in[n];
    }
}

/* calc_Q_matrix (sub_model.cpp:99-114): Q_ij = rates[spec[pos(i,j)]] * (tr ? pi_j : 1),
 * Q_ii = -sum, accumulated by `neg_sum -= value` in column order. */
void orc_build_Q(int S, int tr, const int *spec, const double *rates, const double *freqs_in, double *Q) {
    double pi[64];
    orc_freqs(S, freqs_in, pi);
    for (int i = 0; i < S; ++i) {
        double neg = 0.0;
        for (int j = 0; j < S; ++j) {
            if (i == j) continue;
            double v = rates[spec[orc_rate_pos(S, tr, i, j)]];
            if (tr) v *= pi[j];
            neg -= v;
            Q[i * S + j] = v;
        }
        Q[i * S + i] = neg;
    }
}

/* ------------------------------------------------------------------ */
/* P(t) builders                                                        */
/* ------------------------------------------------------------------ */

enum { ORC_P_REFCOMPAT = 0, ORC_P_EIG = 1, ORC_P_PADE = 2 };

static void matmul(int S, const double *A, const double *B, double *C) {
    /* marth::square_matrix::multiply (marth.cpp:50-61): prod = 0, += a_ik*b_kj, k ascending */
    for (int i = 0; i < S; ++i)
        for (int j = 0; j < S; ++j) {
            double s = 0;
            for (int k = 0; k < S; ++k) s += A[i * S + k] * B[k * S + j];
            C[i * S + j] = s;
        }
}

/* marth::square_matrix::exponential(P, r, k=20) (marth.cpp:129-140) as called
 * by sub_model::set_P_matrix (sub_model.cpp:115-127): P = I + Q*t/2^20, then
 * 20 squarings through power(2) (marth.cpp:106-111). */
static void p_refcompat(int S, const double *Q, double t, double *P) {
    double tmp[64 * 64];
    const double scale = pow(2, 20);
    for (int x = 0; x < S; ++x)
        for (int y = 0; y < S; ++y)
            P[x * S + y] = (x == y ? 1.0 : 0.0) + (Q[x * S + y] * t) / scale;
    for (int i = 0; i < 20; ++i) {
        matmul(S, P, P, tmp);
        memcpy(P, tmp, sizeof(double) * S * S);
    }
}

/* Cyclic Jacobi eigen-decomposition of a symmetric matrix A (S x S).
 * On return w holds eigenvalues and V (row-major) the eigenvectors as columns. */
static void jacobi_sym(int S, double *A, double *w, double *V) {
    for (int i = 0; i < S; ++i) for (int j = 0; j < S; ++j) V[i * S + j] = (i == j);
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0;
        for (int i = 0; i < S; ++i) for (int j = i + 1; j < S; ++j) off += A[i * S + j] * A[i * S + j];
        if (off < 1e-300) break;
        for (int p = 0; p < S; ++p)
            for (int q = p + 1; q < S; ++q) {
                double apq = A[p * S + q];
                if (fabs(apq) < 1e-300) continue;
                double app = A[p * S + p], aqq = A[q * S + q];
                double theta = (aqq - app) / (2 * apq);
                double tt = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1));
                double c = 1 / sqrt(tt * tt + 1), s = tt * c;
                for (int k = 0; k < S; ++k) {
                    double akp = A[k * S + p], akq = A[k * S + q];
                    A[k * S + p] = c * akp - s * akq;
                    A[k * S + q] = s * akp + c * akq;
                }
                for (int k = 0; k < S; ++k) {
                    double apk = A[p * S + k], aqk = A[q * S + k];
                    A[p * S + k] = c * apk - s * aqk;
                    A[q * S + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < S; ++k) {
                    double vkp = V[k * S + p], vkq = V[k * S + q];
                    V[k * S + p] = c * vkp - s * vkq;
                    V[k * S + q] = s * vkp + c * vkq;
                }
            }
    }
    for (int i = 0; i < S; ++i) w[i] = A[i * S + i];
}

/* Reversible Q with stationary pi: B = D^1/2 Q D^-1/2 is symmetric;
 * P(t) = D^-1/2 U exp(L t) U^T D^1/2. */
static void p_eig(int S, const double *Q, const double *pi, double t, double *P) {
    double B[64 * 64], U[64 * 64], w[64], sq[64];
    for (int i = 0; i < S; ++i) sq[i] = sqrt(pi[i]);
    for (int i = 0; i < S; ++i)
        for (int j = 0; j < S; ++j) B[i * S + j] = sq[i] * Q[i * S + j] / sq[j];
    for (int i = 0; i < S; ++i)
        for (int j = i + 1; j < S; ++j) {
            double a = 0.5 * (B[i * S + j] + B[j * S + i]);
            B[i * S + j] = B[j * S + i] = a;
        }
    jacobi_sym(S, B, w, U);
    for (int i = 0; i < S; ++i)
        for (int j = 0; j < S; ++j) {
            double s = 0;
            for (int k = 0; k < S; ++k) s += U[i * S + k] * exp(w[k] * t) * U[j * S + k];
            P[i * S + j] = s * sq[j] / sq[i];
        }
}

/* Gaussian elimination with partial pivoting: solves A X = B (n x n, n x n). */
static void solve(int n, double *A, double *B) {
    for (int c = 0; c < n; ++c) {
        int piv = c;
        for (int r = c + 1; r < n; ++r) if (fabs(A[r * n + c]) > fabs(A[piv * n + c])) piv = r;
        if (piv != c) {
            for (int k = 0; k < n; ++k) {
                double x = A[c * n + k]; A[c * n + k] = A[piv * n + k]; A[piv * n + k] = x;
                x = B[c * n + k]; B[c * n + k] = B[piv * n + k]; B[piv * n + k] = x;
            }
        }
        for (int r = 0; r < n; ++r) {
            if (r == c) continue;
            double f = A[r * n + c] / A[c * n + c];
            if (f == 0) continue;
            for (int k = 0; k < n; ++k) { A[r * n + k] -= f * A[c * n + k]; B[r * n + k] -= f * B[c * n + k]; }
        }
    }
    for (int r = 0; r < n; ++r) { double d = A[r * n + r]; for (int k = 0; k < n; ++k) B[r * n + k] /= d; }
}

/* Scaling and squaring with the [13/13] Pade approximant (Higham, SIAM J.
 * Matrix Anal. Appl. 26 (2005) 1179-1193, Algorithm 2.3). */
static void p_pade(int S, const double *Q, double t, double *P) {
    static const double b[14] = {64764752532480000., 32382376266240000., 7771770303897600.,
        1187353796428800., 129060195264000., 10559470521600., 670442572800., 33522128640.,
        1323241920., 40840800., 960960., 16380., 182., 1.};
    int n = S, nn = S * S;
    double *A = malloc(sizeof(double) * nn * 12);
    double *A2 = A + nn, *A4 = A2 + nn, *A6 = A4 + nn, *U = A6 + nn, *V = U + nn, *T = V + nn,
           *N = T + nn, *D = N + nn, *W = D + nn;
    double norm = 0;
    for (int i = 0; i < n; ++i) {
        double s = 0;
        for (int j = 0; j < n; ++j) s += fabs(Q[i * n + j] * t);
        if (s > norm) norm = s;
    }
    int sq = 0;
    const double theta13 = 5.371920351148152;
    if (norm > theta13) sq = (int)ceil(log2(norm / theta13));
    double sc = ldexp(1.0, -sq);
    for (int i = 0; i < nn; ++i) A[i] = Q[i] * t * sc;
    matmul(n, A, A, A2); matmul(n, A2, A2, A4); matmul(n, A4, A2, A6);
    for (int i = 0; i < nn; ++i) T[i] = b[13] * A6[i] + b[11] * A4[i] + b[9] * A2[i];
    matmul(n, A6, T, W);
    for (int i = 0; i < nn; ++i) W[i] += b[7] * A6[i] + b[5] * A4[i] + b[3] * A2[i];
    for (int i = 0; i < n; ++i) W[i * n + i] += b[1];
    matmul(n, A, W, U);
    for (int i = 0; i < nn; ++i) T[i] = b[12] * A6[i] + b[10] * A4[i] + b[8] * A2[i];
    matmul(n, A6, T, V);
    for (int i = 0; i < nn; ++i) V[i] += b[6] * A6[i] + b[4] * A4[i] + b[2] * A2[i];
    for (int i = 0; i < n; ++i) V[i * n + i] += b[0];
    for (int i = 0; i < nn; ++i) { N[i] = V[i] + U[i]; D[i] = V[i] - U[i]; }
    solve(n, D, N);
    for (int s = 0; s < sq; ++s) { matmul(n, N, N, T); memcpy(N, T, sizeof(double) * nn); }
    memcpy(P, N, sizeof(double) * nn);
    free(A);
}

int orc_pmatrix(int S, const double *Q, const double *pi, double t, int builder, double *P) {
    if (S < 1 || S > 64) return ORC_ERR;
    if (builder == ORC_P_REFCOMPAT) p_refcompat(S, Q, t, P);
    else if (builder == ORC_P_EIG) p_eig(S, Q, pi, t, P);
    else if (builder == ORC_P_PADE) p_pade(S, Q, t, P);
    else return ORC_ERR;
    return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* rate_model (rate_model.cpp:103-127, rate_model.h:64-67)             */
/* ------------------------------------------------------------------ */

/* get_time_mod_branch_length with the reference's float distance_to_root and
 * float cut-offs / rates (rate_model.h:78-80).  cut[i] are the sorted
 * cut-off times, cut_rate[i] = rates_in_time[rate_class of cut-off i]. */
double orc_time_mod_length(double branch_length, float distance_to_root,
                           int n_cut, const float *cut, const float *cut_rate) {
    const double branch_end = distance_to_root + branch_length;
    double nb = 0.0;
    double prev = -0.0;
    for (int i = 0; i < n_cut; ++i) {
        if (distance_to_root < cut[i] && branch_end > prev) {
            if (distance_to_root >= prev && branch_end <= cut[i])
                nb += (branch_end - distance_to_root) * cut_rate[i];
            else if (distance_to_root < prev && branch_end <= cut[i])
                nb += (branch_end - prev) * cut_rate[i];
            else if (distance_to_root >= prev && branch_end > cut[i])
                nb += (cut[i] - distance_to_root) * cut_rate[i];
            else if (distance_to_root < prev && branch_end > cut[i])
                nb += (cut[i] - prev) * cut_rate[i];
        }
        prev = cut[i];
    }
    if (distance_to_root < prev && branch_end > prev) nb += branch_end - prev;
    else if (distance_to_root > prev) nb = branch_length;
    return nb;
}

/* Effective length of the edge above each node, as simpleML::calculate_likelihood
 * computes it (simpleML.cpp:127-184): the child edge of parent p uses
 * get_time_mod_branch_length(b_child, dist_p) * rate_clade(clade_no_p), the
 * clade number switches at nodes listed in clade_of_node (tree.h:288-292),
 * distances accumulate raw lengths (simpleML.cpp:134,138), root dist = 0.
 * With no cut-offs this zeroes the root's child edges (rate_model.cpp:106,122)
 * unless root_quirk == 0, in which case raw lengths are used at the root. */
static void eff_rec(int v, int clade_no, double dist, const int *left, const int *right, const double *blen,
                    int n_cut, const float *cut, const float *cut_rate, const int *clade_of_node,
                    int n_clade, const float *clade_rate, int root, int root_quirk, double *eff) {
    if (left[v] < 0 && right[v] < 0) return;
    if (clade_of_node && clade_of_node[v] >= 0) clade_no = clade_of_node[v];
    double rate = (clade_no < n_clade) ? (double)clade_rate[clade_no] : 1.0;
    int ch[2] = {left[v], right[v]};
    for (int k = 0; k < 2; ++k) {
        int c = ch[k];
        if (c < 0) continue;
        double len;
        if (v == root && !root_quirk && n_cut == 0) len = blen[c];
        else len = orc_time_mod_length(blen[c], (float)dist, n_cut, cut, cut_rate);
        eff[c] = len * rate;
        eff_rec(c, clade_no, dist + blen[c], left, right, blen, n_cut, cut, cut_rate, clade_of_node,
                n_clade, clade_rate, root, root_quirk, eff);
    }
}
void orc_effective_lengths(int n_nodes, int root, const int *left, const int *right, const double *blen,
                           int n_cut, const float *cut, const float *cut_rate,
                           const int *clade_of_node, int n_clade, const float *clade_rate,
                           int root_quirk, double *eff) {
    for (int i = 0; i < n_nodes; ++i) eff[i] = 0.0;
    eff_rec(root, 0, 0.0, left, right, blen, n_cut, cut, cut_rate, clade_of_node, n_clade, clade_rate,
            root, root_quirk, eff);
}

/* ------------------------------------------------------------------ */
/* Discrete gamma rate categories (Yang 1994, J Mol Evol 39:306-314)    */
/* ------------------------------------------------------------------ */

/* Regularised lower incomplete gamma P(a, x): series for x < a+1, Lentz
 * continued fraction for Q otherwise (Press et al., Numerical Recipes 6.2). */
double orc_gammainc(double a, double x) {
    if (x <= 0) return 0.0;
    double gln = lgamma(a);
    if (x < a + 1) {
        double ap = a, sum = 1.0 / a, del = sum;
        for (int n = 0; n < 10000; ++n) {
            ap += 1; del *= x / ap; sum += del;
            if (fabs(del) < fabs(sum) * 1e-17) break;
        }
        return sum * exp(-x + a * log(x) - gln);
    }
    double b = x + 1 - a, c = 1 / 1e-300, d = 1 / b, h = d;
    for (int i = 1; i < 10000; ++i) {
        double an = -i * (i - a);
        b += 2;
        d = an * d + b; if (fabs(d) < 1e-300) d = 1e-300;
        c = b + an / c; if (fabs(c) < 1e-300) c = 1e-300;
        d = 1 / d;
        double del = d * c;
        h *= del;
        if (fabs(del - 1) < 1e-17) break;
    }
    return 1.0 - exp(-x + a * log(x) - gln) * h;
}

/* Inverse of P(a, .) by bracketing bisection refined with Newton steps. */
double orc_gammaincinv(double a, double p) {
    if (p <= 0) return 0.0;
    if (p >= 1) return INFINITY;
    double lo = 0, hi = 1;
    while (orc_gammainc(a, hi) < p) { lo = hi; hi *= 2; if (hi > 1e300) return INFINITY; }
    double x = 0.5 * (lo + hi);
    for (int it = 0; it < 400; ++it) {
        double f = orc_gammainc(a, x) - p;
        if (f > 0) hi = x; else lo = x;
        double dens = exp(-x + (a - 1) * log(x) - lgamma(a));
        double xn = x - f / dens;
        if (!(xn > lo && xn < hi)) xn = 0.5 * (lo + hi);
        if (fabs(xn - x) <= 1e-16 * fabs(x) || hi - lo <= 1e-300) { x = xn; break; }
        x = xn;
    }
    return x;
}

/* Mean rate of each of C equal-probability categories of Gamma(alpha, alpha). */
void orc_gamma_rates(double alpha, int C, double *rates) {
    double prev_I = 0.0;
    for (int k = 0; k < C; ++k) {
        double I;
        if (k == C - 1) I = 1.0;
        else {
            double q = orc_gammaincinv(alpha, (double)(k + 1) / C) / alpha;
            I = orc_gammainc(alpha + 1, q * alpha);
        }
        rates[k] = (I - prev_I) * C;
        prev_I = I;
    }
}

/* ------------------------------------------------------------------ */
/* Pruning likelihood (simpleML.cpp:127-197, simpleML.h:123-141)        */
/* ------------------------------------------------------------------ */

typedef struct {
    int n_nodes, root;
    const int *left, *right;
    const double *eff;         /* effective length of the edge above node */
    int m, S, C;
    const uint32_t *masks;     /* [node][m], read for tips only */
    const double *cat_rate, *cat_w;
    const double *Q, *pi;      /* pi used by the EIG builder */
    int builder, per_site_p;
    const double *root_w;      /* S weights at the root: pi (TR) or ones (non-TR) */
} orc_problem;

static int *postorder(int n_nodes, int root, const int *left, const int *right, int *n_out) {
    /* recursion order of calculate_likelihood: left subtree, right subtree, node */
    int *order = malloc(sizeof(int) * n_nodes), *stack = malloc(sizeof(int) * 2 * n_nodes);
    char *done = calloc(n_nodes, 1);
    int sp = 0, k = 0;
    stack[sp++] = root;
    while (sp) {
        int v = stack[sp - 1];
        if (!done[v]) {
            done[v] = 1;
            if (right[v] >= 0) stack[sp++] = right[v];
            if (left[v] >= 0) stack[sp++] = left[v];
        } else {
            --sp;
            order[k++] = v;
        }
    }
    free(stack); free(done);
    *n_out = k;
    return order;
}

/* branch_likelihood (simpleML.cpp:186-197): out_i (=|*=) sum_j P_ij in_j,
 * sum from 0.0 in ascending j. */
static void branch(int S, const double *P, const double *in, double *out, int multiply) {
    for (int i = 0; i < S; ++i) {
        double pi = 0.0;
        for (int j = 0; j < S; ++j) pi += in[j] * P[i * S + j];
        if (multiply) out[i] *= pi; else out[i] = pi;
    }
}

/* Computes per-node CLVs ([node][site][cat][S], unscaled values times
 * 2^-scale[node][site]) and per-site lnL.  clv/scale may be NULL when not wanted. */
static int run_pruning(const orc_problem *p, double *clv_out, int *scale_out, double *site_lnl, double *total) {
    int n = p->n_nodes, m = p->m, S = p->S, C = p->C;
    size_t per = (size_t)C * S;
    double *clv = clv_out ? clv_out : malloc(sizeof(double) * (size_t)n * m * per);
    int *scl = scale_out ? scale_out : calloc((size_t)n * m, sizeof(int));
    if (!clv || !scl) return ORC_ERR;
    memset(scl, 0, sizeof(int) * (size_t)n * m);
    int n_order;
    int *order = postorder(n, p->root, p->left, p->right, &n_order);
    double *P = malloc(sizeof(double) * (size_t)n * C * S * S);
    double *Ps = malloc(sizeof(double) * S * S * 2);
    double *tipv = malloc(sizeof(double) * S);
    /* P per (edge, category), hoisted unless per_site_p (the reference recomputes
     * P for every site and edge, simpleML.cpp:187). */
    if (!p->per_site_p)
        for (int v = 0; v < n; ++v)
            if (v != p->root)
                for (int c = 0; c < C; ++c)
                    orc_pmatrix(S, p->Q, p->pi, p->eff[v] * p->cat_rate[c], p->builder, P + ((size_t)v * C + c) * S * S);
    for (int oi = 0; oi < n_order; ++oi) {
        int v = order[oi];
        int L = p->left[v], R = p->right[v];
        if (L < 0 && R < 0) {          /* tip: 1.0 for every set state (simpleML.h:41-57) */
            for (int s = 0; s < m; ++s) {
                uint32_t mk = p->masks[(size_t)v * m + s];
                for (int c = 0; c < C; ++c)
                    for (int i = 0; i < S; ++i)
                        clv[((size_t)v * m + s) * per + c * S + i] = (mk >> i) & 1u ? 1.0 : 0.0;
            }
            continue;
        }
        for (int s = 0; s < m; ++s) {
            double *out = clv + ((size_t)v * m + s) * per;
            int sc = 0;
            for (int c = 0; c < C; ++c) {
                const double *PL = NULL, *PR = NULL;
                if (p->per_site_p) {
                    if (L >= 0) { orc_pmatrix(S, p->Q, p->pi, p->eff[L] * p->cat_rate[c], p->builder, Ps); PL = Ps; }
                    if (R >= 0) { orc_pmatrix(S, p->Q, p->pi, p->eff[R] * p->cat_rate[c], p->builder, Ps + S * S); PR = Ps + S * S; }
                } else {
                    if (L >= 0) PL = P + ((size_t)L * C + c) * S * S;
                    if (R >= 0) PR = P + ((size_t)R * C + c) * S * S;
                }
                if (L >= 0) branch(S, PL, clv + ((size_t)L * m + s) * per + c * S, out + c * S, 0);
                if (R >= 0) branch(S, PR, clv + ((size_t)R * m + s) * per + c * S, out + c * S, L >= 0);
            }
            if (L >= 0) sc += scl[(size_t)L * m + s];
            if (R >= 0) sc += scl[(size_t)R * m + s];
            /* exact power-of-two rescale, only when the values approach underflow */
            double mx = 0;
            for (size_t k = 0; k < per; ++k) if (out[k] > mx) mx = out[k];
            if (mx > 0 && mx < 0x1p-500) {
                int e; frexp(mx, &e);
                for (size_t k = 0; k < per; ++k) out[k] = ldexp(out[k], -e);
                sc += e;
            }
            scl[(size_t)v * m + s] = sc;
        }
    }
    /* get_log_likelihood_from_root (simpleML.h:123-141), extended with category
     * weights; the total is the plain running sum of site logs. */
    double tot = 0.0;
    int r = p->root;
    for (int s = 0; s < m; ++s) {
        const double *x = clv + ((size_t)r * m + s) * per;
        double like = 0.0;
        if (C == 1) {
            for (int i = 0; i < S; ++i) like += x[i] * p->root_w[i];
        } else {
            for (int c = 0; c < C; ++c) {
                double lc = 0.0;
                for (int i = 0; i < S; ++i) lc += x[c * S + i] * p->root_w[i];
                like += p->cat_w[c] * lc;
            }
        }
        double l = log(like) + scl[(size_t)r * m + s] * 0.6931471805599453094;
        if (site_lnl) site_lnl[s] = l;
        tot += l;
    }
    if (total) *total = tot;
    free(order); free(P); free(Ps); free(tipv);
    if (!clv_out) free(clv);
    if (!scale_out) free(scl);
    return ORC_OK;
}

int orc_loglik(int n_nodes, int root, const int *left, const int *right, const double *eff,
               int m, int S, const uint32_t *masks, int C, const double *cat_rate, const double *cat_w,
               const double *Q, const double *pi, int builder, int per_site_p, const double *root_w,
               double *site_lnl, double *total) {
    orc_problem p = {n_nodes, root, left, right, eff, m, S, C, masks, cat_rate, cat_w, Q, pi,
                     builder, per_site_p, root_w};
    return run_pruning(&p, NULL, NULL, site_lnl, total);
}

/* Per-node CLVs (and scalers) for every node, for checking materialised partials. */
int orc_node_clvs(int n_nodes, int root, const int *left, const int *right, const double *eff,
                  int m, int S, const uint32_t *masks, int C, const double *cat_rate, const double *cat_w,
                  const double *Q, const double *pi, int builder, const double *root_w,
                  double *clv, int *scale, double *site_lnl, double *total) {
    orc_problem p = {n_nodes, root, left, right, eff, m, S, C, masks, cat_rate, cat_w, Q, pi,
                     builder, 0, root_w};
    return run_pruning(&p, clv, scale, site_lnl, total);
}

/* draw_normalized_likelihood_on_nodes (simpleML.cpp:199-231): for each internal
 * node, per site, L_i / sum where `sum` keeps accumulating across sites
 * (it is not reset per site, :204, :211).  Category CLVs are first mixed with
 * the category weights.  out: [node][m][S] (tips left untouched). */
int orc_node_partials(int n_nodes, int root, const int *left, const int *right, const double *eff,
                      int m, int S, const uint32_t *masks, int C, const double *cat_rate, const double *cat_w,
                      const double *Q, const double *pi, int builder, const double *root_w,
                      int reset_sum_per_site, double *out) {
    size_t per = (size_t)C * S;
    double *clv = malloc(sizeof(double) * (size_t)n_nodes * m * per);
    int *scl = malloc(sizeof(int) * (size_t)n_nodes * m);
    orc_problem p = {n_nodes, root, left, right, eff, m, S, C, masks, cat_rate, cat_w, Q, pi, builder, 0, root_w};
    int rc = run_pruning(&p, clv, scl, NULL, NULL);
    if (rc) { free(clv); free(scl); return rc; }
    for (int v = 0; v < n_nodes; ++v) {
        if (left[v] < 0 && right[v] < 0) continue;
        double sum = 0.0;
        for (int s = 0; s < m; ++s) {
            double x[64];
            const double *src = clv + ((size_t)v * m + s) * per;
            double f = ldexp(1.0, scl[(size_t)v * m + s]);
            for (int i = 0; i < S; ++i) {
                double a = 0;
                if (C == 1) a = src[i];
                else for (int c = 0; c < C; ++c) a += cat_w[c] * src[c * S + i];
                x[i] = a * f;
            }
            if (reset_sum_per_site) sum = 0.0;
            for (int i = 0; i < S; ++i) sum += x[i];
            for (int i = 0; i < S; ++i) out[((size_t)v * m + s) * S + i] = x[i] / sum;
        }
    }
    free(clv); free(scl);
    return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* NNI neighbourhood as treebender --nni all emits it                   */
/* (treebender.cpp:897-915, tree.cpp:2468-2548)                         */
/* ------------------------------------------------------------------ */

static int n_sub_tips(const orc_tree *t, int v) {
    if (v < 0) return 0;
    if (t->left[v] < 0 && t->right[v] < 0) return 1;
    return n_sub_tips(t, t->left[v]) + n_sub_tips(t, t->right[v]);
}

/* get_internal_node_by_number (tree.cpp:2455-2464): preorder count of internal nodes */
static int internal_by_number(const orc_tree *t, int v, int *found, int no) {
    if (v < 0 || (t->left[v] < 0 && t->right[v] < 0)) return no;
    --no;
    if (no == 0) { *found = v; return no; }
    if (t->left[v] >= 0) no = internal_by_number(t, t->left[v], found, no);
    if (*found < 0 && t->right[v] >= 0) no = internal_by_number(t, t->right[v], found, no);
    return no;
}

/* get_internal_branch_by_number_unrooted (tree.cpp:2466-2478) */
static int branch_by_number(const orc_tree *t, int no) {
    int found = -1;
    if (no == 0) return -1;
    int r = t->root;
    if (n_sub_tips(t, t->left[r]) < 2) internal_by_number(t, t->right[r], &found, no);
    else if (n_sub_tips(t, t->right[r]) < 2) internal_by_number(t, t->left[r], &found, no);
    else internal_by_number(t, r, &found, no + 1);
    return found;
}

/* tree::interchange (tree.cpp:2480-2518) on arrays */
static void interchange(orc_tree *t, int leaf, int use_left) {
    if (leaf < 0 || t->parent[leaf] < 0 || (t->left[leaf] < 0 && t->right[leaf] < 0)) return;
    int par = t->parent[leaf], partner = -1, r = t->root;
    if (t->left[par] >= 0 && t->left[par] != leaf) {
        if (par != r) partner = t->left[par];
        else if (t->left[t->left[par]] >= 0) partner = t->left[t->left[par]];
    } else if (t->right[par] >= 0 && t->right[par] != leaf) {
        if (par != r) partner = t->right[par];
        else if (t->left[t->right[par]] >= 0) partner = t->left[t->right[par]];
    }
    if (partner < 0) return;
    int tmp = -1;
    if (use_left && t->left[leaf] >= 0) { tmp = t->left[leaf]; t->left[leaf] = partner; }
    if (!use_left && t->right[leaf] >= 0) { tmp = t->right[leaf]; t->right[leaf] = partner; }
    if (tmp >= 0) {
        int pp = t->parent[partner];
        t->parent[tmp] = pp;
        if (t->left[pp] == partner) t->left[pp] = tmp;
        else if (t->right[pp] == partner) t->right[pp] = tmp;
        t->parent[partner] = leaf;
    }
}

static orc_tree *tree_copy(const orc_tree *a) {
    orc_tree *t = calloc(1, sizeof(orc_tree));
    for (int i = 0; i < a->n; ++i) tree_new_node(t);
    t->root = a->root;
    for (int i = 0; i < a->n; ++i) {
        t->left[i] = a->left[i]; t->right[i] = a->right[i]; t->parent[i] = a->parent[i];
        t->blen[i] = a->blen[i]; t->label[i] = a->label[i] ? strdup(a->label[i]) : NULL;
    }
    return t;
}

/* Writes the newick texts of the original tree followed by the 2*(n-4) NNI
 * variants (branches 2..n-3, left then right swap) one per line into out.
 * Returns the number of trees, or -1.  Needs cap large enough. */
int orc_nni_reference(const char *newick, char *out, int cap) {
    orc_tree *t = orc_tree_parse(newick);
    if (!t || t->n == 0) { orc_tree_free(t); return -1; }
    int ntips = n_sub_tips(t, t->root);
    int used = 0, count = 0;
    char *buf = malloc(1 << 20);
    int len = orc_tree_newick(t, buf, 1 << 20);
    if (used + len + 1 < cap) { memcpy(out + used, buf, len); used += len; out[used++] = '\n'; }
    ++count;
    if (ntips >= 4) {
        for (int b = 2; b <= ntips - 3; ++b) {
            for (int side = 0; side < 2; ++side) {
                orc_tree *c = tree_copy(t);
                int v = branch_by_number(c, b);
                if (v >= 0) interchange(c, v, side == 0);
                len = orc_tree_newick(c, buf, 1 << 20);
                if (used + len + 1 < cap) { memcpy(out + used, buf, len); used += len; out[used++] = '\n'; }
                ++count;
                orc_tree_free(c);
            }
        }
    }
    if (used < cap) out[used] = 0;
    free(buf);
    orc_tree_free(t);
    return count;
}

static orc_tree *tree_from_arrays(int n_nodes, int root, const int *left, const int *right) {
    orc_tree *t = calloc(1, sizeof(orc_tree));
    for (int i = 0; i < n_nodes; ++i) tree_new_node(t);
    t->root = root;
    for (int i = 0; i < n_nodes; ++i) { t->left[i] = left[i]; t->right[i] = right[i]; }
    for (int i = 0; i < n_nodes; ++i) {
        if (left[i] >= 0) t->parent[left[i]] = i;
        if (right[i] >= 0) t->parent[right[i]] = i;
    }
    return t;
}

/* NNI variants of an array tree as (node, side) pairs, side 0 = left swap.
 * mode 0: treebender --nni all (branches 2..n-3 through
 * get_internal_branch_by_number_unrooted, treebender.cpp:897-915);
 * mode 1: every internal edge of the unrooted tree (both root children
 * count as one edge), 2*(n-3) variants.  Returns the count. */
int orc_nni_list(int n_nodes, int root, const int *left, const int *right, int mode, int *out_node, int *out_side) {
    orc_tree *t = tree_from_arrays(n_nodes, root, left, right);
    int ntips = n_sub_tips(t, root), k = 0;
    if (mode == 0) {
        if (ntips >= 4)
            for (int b = 2; b <= ntips - 3; ++b) {
                int v = branch_by_number(t, b);
                for (int side = 0; side < 2; ++side) { out_node[k] = v; out_side[k] = side; ++k; }
            }
    } else {
        int a = left[root], bb = right[root];
        int a_int = a >= 0 && left[a] >= 0, b_int = bb >= 0 && left[bb] >= 0;
        for (int v = 0; v < n_nodes; ++v) {
            if (v == root || left[v] < 0) continue;
            if (t->parent[v] == root) {
                if (!(a_int && b_int)) continue;      /* root edge is not internal */
                if (v == bb) continue;                /* count the root edge once */
            }
            for (int side = 0; side < 2; ++side) { out_node[k] = v; out_side[k] = side; ++k; }
        }
    }
    orc_tree_free(t);
    return k;
}

/* Applies tree::interchange(node, side == 0) (tree.cpp:2480-2518) to copies
 * of the arrays. */
void orc_nni_apply(int n_nodes, int root, const int *left, const int *right, int node, int side,
                   int *out_left, int *out_right) {
    orc_tree *t = tree_from_arrays(n_nodes, root, left, right);
    if (node >= 0) interchange(t, node, side == 0);
    for (int i = 0; i < n_nodes; ++i) { out_left[i] = t->left[i]; out_right[i] = t->right[i]; }
    orc_tree_free(t);
}

/* ------------------------------------------------------------------ */
/* Forward simulation (simpleML.cpp:233-292), seeded                    */
/* ------------------------------------------------------------------ */

static uint64_t sm64(uint64_t *x) {
    uint64_t z = (*x += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
static double urand(uint64_t *x) { return (sm64(x) >> 11) * 0x1.0p-53; }

/* Root state from freqs, then per edge inverse-CDF sampling on row P[anc]
 * with P(raw branch length) (simpleML.cpp:274-283), per-site rate category
 * drawn uniformly.  states: [node][m]. */
int orc_simulate(int n_nodes, int root, const int *left, const int *right, const double *blen,
                 int m, int S, const double *Q, const double *pi, const double *freqs,
                 int C, const double *cat_rate, uint64_t seed, int *states) {
    int n_order;
    int *order = postorder(n_nodes, root, left, right, &n_order);
    double *P = malloc(sizeof(double) * (size_t)n_nodes * C * S * S);
    for (int v = 0; v < n_nodes; ++v)
        for (int c = 0; c < C; ++c)
            orc_pmatrix(S, Q, pi, blen[v] * cat_rate[c], ORC_P_PADE, P + ((size_t)v * C + c) * S * S);
    uint64_t st = seed;
    for (int s = 0; s < m; ++s) {
        int c = C > 1 ? (int)(urand(&st) * C) : 0;
        double u = urand(&st);
        int rs = S - 1;
        for (int j = 0; j < S; ++j) { u -= freqs[j]; if (u <= 0) { rs = j; break; } }
        states[(size_t)root * m + s] = rs;
        for (int oi = n_order - 1; oi >= 0; --oi) {      /* reverse postorder = parents first */
            int v = order[oi];
            int ch[2] = {left[v], right[v]};
            for (int k = 0; k < 2; ++k) {
                int w = ch[k];
                if (w < 0) continue;
                int anc = states[(size_t)v * m + s];
                const double *row = P + ((size_t)w * C + c) * S * S + anc * S;
                double r = urand(&st);
                int tr = anc;
                for (int j = 0; j < S; ++j) { r -= row[j]; if (r <= 0) { tr = j; break; } }
                states[(size_t)w * m + s] = tr;
            }
        }
    }
    free(order); free(P);
    return ORC_OK;
}

/* The reference's cost model as a CPU baseline: per-site recursion with P
 * rebuilt by refcompat squaring for every (site, edge), as simpleML does.
 * Returns the wall time of the likelihood call (seconds) via *seconds. */
#include <time.h>
int orc_time_reference_cost(int n_nodes, int root, const int *left, const int *right, const double *eff,
                            int m, int S, const uint32_t *masks, int C, const double *cat_rate,
                            const double *cat_w, const double *Q, const double *root_w,
                            double *total, double *seconds) {
    struct timespec a, b;
    clock_gettime(CLOCK_MONOTONIC, &a);
    int rc = orc_loglik(n_nodes, root, left, right, eff, m, S, masks, C, cat_rate, cat_w, Q, root_w,
                        ORC_P_REFCOMPAT, 1, root_w, NULL, total);
    clock_gettime(CLOCK_MONOTONIC, &b);
    *seconds = (b.tv_sec - a.tv_sec) + 1e-9 * (b.tv_nsec - a.tv_nsec);
    return rc;
}
