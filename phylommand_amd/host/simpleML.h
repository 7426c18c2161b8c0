// simpleML.h -- phylommand's simpleML likelihood class (simpleML.h:30-142)
// with the compute done by libphylo_hip on the GPU.
//
// Kept from the reference: the call surface (init/reset, set_char per taxon
// and character, calculate_log_likelihood with and without a rate_model and
// its two aliases, draw_normalized_likelihood_on_nodes, simulate_chars,
// get_n_char), the effective branch lengths (time-rate integral x clade rate
// of the parent's clade, distances accumulated from raw lengths, so the
// root's child edges get length 0 with no time cut-offs;
// simpleML.cpp:127-184), the unweighted root sum of non-reversible models
// (simpleML.h:132-133), and the return value 0.0 with a message when a tip
// lacks a state (check_nodes, simpleML.cpp:52-71).
// Added: rate categories, per-site log-likelihoods, NNI neighbourhood
// scores, a choice of P(t) builder and of FP32/FP64 partial storage.
#ifndef PHYLO_AMD_SIMPLEML_H
#define PHYLO_AMD_SIMPLEML_H

#include <cstdint>
#include <string>
#include <unordered_map>
#include <vector>

#include "io.h"
#include "rate_model.h"
#include "sub_model.h"
#include "phylo_tree.h"

struct phylo_ctx;

namespace phylo_amd {

// Simulated characters: per taxon its name and one state set per character.
// Converts to a std::vector of any character-vector type with set_taxon and
// add_character -- this library's character_vector, or the caller's own
// (treeator assigns the result to its vector<character_vector>,
// treeator.cpp:735).
struct simulated_rows {
    std::vector<std::string> taxa;
    std::vector<std::vector<state_set>> rows;
    template <class CV>
    operator std::vector<CV>() const {
        std::vector<CV> out(taxa.size());
        for (size_t i = 0; i < taxa.size(); ++i) {
            out[i].set_taxon(taxa[i]);
            for (const state_set& x : rows[i]) out[i].add_character(x);
        }
        return out;
    }
};

class simpleML : public tree {
  public:
    // The engine options start from the environment, so a program built
    // against these classes unchanged (treeator) can choose them:
    // PHYLO_AMD_P_BUILDER = refcompat | eigen | pade (default: eigen for
    // reversible models, pade otherwise), PHYLO_AMD_PRECISION = fp64 |
    // fp32 (default fp32) and PHYLO_AMD_SITE_PATTERNS = 1 (default 0).
    // refcompat with fp64 reproduces the reference's own values to about
    // 1e-11.
    simpleML();
    ~simpleML() override;
    simpleML(const simpleML&) = delete;
    simpleML& operator=(const simpleML&) = delete;

    void init();
    void reset() { init(); }
    void set_char(const int char_no, const std::string& taxon, const state_set& state, sub_model& model);
    double calculate_log_likelihood(sub_model& model);
    double calculate_log_likelihood(sub_model& model, rate_model& rate_mod);
    double calculate_likelihood_rate_change_at_nodes(rate_model& rate_mod, sub_model& model) {
        return calculate_log_likelihood(model, rate_mod);
    }
    double calculate_likelihood_rate_change_in_time(rate_model& rate_mod, sub_model& model) {
        return calculate_log_likelihood(model, rate_mod);
    }
    simulated_rows simulate_chars(const std::vector<double>& charfreq, const unsigned int n_char, sub_model& model);
    // Appends "[&P={...},P={...}]" (per site, L/sum with the running sum of the
    // reference, simpleML.cpp:199-231) to every internal node's label.
    void draw_normalized_likelihood_on_nodes();
    unsigned int get_n_char() const;

    // ---- engine options and extra results ----
    enum { P_REFCOMPAT = 0, P_EIGEN = 1, P_PADE = 2, P_AUTO = -1 };
    void set_rate_categories(const std::vector<double>& rates, const std::vector<double>& weights);
    void set_precision_fp64(bool fp64) { fp64_ = fp64; }
    void set_p_builder(int builder) { builder_ = builder; }
    // false: the root's child edges keep their lengths (textbook likelihood)
    void set_root_edge_compat(bool on) { root_compat_ = on; }
    void set_simulation_seed(uint64_t seed) { seed_ = seed; seeded_ = true; }
    // Site-pattern compression (sketched in the reference's MLtree.h:44-47):
    // identical columns of the tip data are evaluated once, weighted by
    // their count; per-site results are expanded back to every site.
    void set_site_patterns(bool on) { patterns_ = on; }
    unsigned int n_patterns() const { return n_patterns_; }
    const std::vector<double>& site_log_likelihoods() const { return site_lnl_; }
    // Scores the NNI neighbourhood (mode 0: treebender --nni all order,
    // mode 1: every internal edge) of the current tree with the model.
    std::vector<double> nni_neighbourhood(sub_model& model, int mode, std::vector<int>* node = nullptr,
                                          std::vector<int>* side = nullptr);
    const std::string& last_error() const { return err_; }

  private:
    bool prepare(sub_model& model, rate_model& rate_mod, std::vector<int>& lib_of_node);
    void effective_lengths(int v, unsigned int clade_no, double dist, rate_model& rate_mod,
                           std::vector<double>& eff) const;
    int tip_of(const std::string& taxon);
    std::vector<std::vector<state_set>> tipdata_;   // per node, per character
    std::unordered_map<std::string, int> tip_of_taxon_;   // find_taxon_tip, cached
    std::vector<double> site_lnl_;
    std::vector<double> cat_rates_{1.0}, cat_w_{1.0};
    std::vector<int> lib_of_node_;
    phylo_ctx* ctx_ = nullptr;
    std::vector<unsigned int> site_pattern_;       // site -> evaluated column (patterns on)
    unsigned int n_patterns_ = 0;
    bool fp64_ = false, root_compat_ = true, seeded_ = false, patterns_ = false;
    int builder_ = P_AUTO;
    unsigned int n_states_ = 0;
    uint64_t seed_ = 0;
    std::string err_;
};

}  // namespace phylo_amd

#ifdef PHYLO_AMD_AS_REFERENCE
// Building phylommand's own sources against these classes (the treeator
// binding): the unqualified name they use.
using phylo_amd::simpleML;
#endif

#endif
