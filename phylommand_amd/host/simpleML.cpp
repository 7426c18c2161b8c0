// simpleML.cpp -- see simpleML.h.
#include "simpleML.h"

#include <cmath>
#include <cstdlib>
#include <iostream>
#include <map>
#include <random>
#include <sstream>

#include "phylo_hip.h"

namespace phylo_amd {

simpleML::simpleML() {
    const char* b = std::getenv("PHYLO_AMD_P_BUILDER");
    if (b) {
        const std::string v(b);
        if (v == "refcompat") builder_ = P_REFCOMPAT;
        else if (v == "eigen") builder_ = P_EIGEN;
        else if (v == "pade") builder_ = P_PADE;
    }
    const char* p = std::getenv("PHYLO_AMD_PRECISION");
    if (p) fp64_ = std::string(p) == "fp64";
    const char* sp = std::getenv("PHYLO_AMD_SITE_PATTERNS");
    if (sp) patterns_ = std::string(sp) == "1";
}

simpleML::~simpleML() {
    if (ctx_) phylo_destroy(ctx_);
}

void simpleML::init() {
    tipdata_.assign(nodes_.size(), {});
    site_lnl_.clear();
    tip_of_taxon_.clear();
}

// find_taxon_tip (a scan over the tips, tree.cpp:1117-1129 in the reference)
// through a hash map, so filling m characters of n taxa costs O(n m), not
// O(n^2 m).  A hit is checked against the tree, a miss rebuilds the map once,
// so the map can never answer differently from the scan.
int simpleML::tip_of(const std::string& taxon) {
    auto valid = [&](int v) {
        return v >= 0 && (size_t)v < nodes_.size() && is_tip(v) && nodes_[v].has_label && nodes_[v].label == taxon;
    };
    auto it = tip_of_taxon_.find(taxon);
    if (it != tip_of_taxon_.end() && valid(it->second)) return it->second;
    tip_of_taxon_.clear();
    for (int v : tips())
        if (nodes_[v].has_label) tip_of_taxon_.emplace(nodes_[v].label, v);   // first tip of a label wins
    it = tip_of_taxon_.find(taxon);
    return it != tip_of_taxon_.end() ? it->second : -1;
}

void simpleML::set_char(const int char_no, const std::string& taxon, const state_set& state, sub_model& model) {
    const int leaf = tip_of(taxon);
    if (leaf < 0 || char_no < 0) return;           // taxa absent from the tree are ignored (simpleML.h:45-46)
    if (tipdata_.size() != nodes_.size()) tipdata_.resize(nodes_.size());
    auto& row = tipdata_[leaf];
    if ((size_t)char_no >= row.size()) row.resize(char_no + 1);
    state_set s;
    for (unsigned int i = 0; i < model.get_n_states() && i < SIZE; ++i) s[i] = state[i];
    row[char_no] = s;
}

unsigned int simpleML::get_n_char() const {
    size_t n = 0;
    for (int v : tips())
        if ((size_t)v < tipdata_.size() && tipdata_[v].size() > n) n = tipdata_[v].size();
    return (unsigned int)n;
}

void simpleML::set_rate_categories(const std::vector<double>& rates, const std::vector<double>& weights) {
    cat_rates_ = rates;
    cat_w_ = weights;
}

// simpleML::calculate_likelihood's branch lengths (simpleML.cpp:127-184).
void simpleML::effective_lengths(int v, unsigned int clade_no, double dist, rate_model& rate_mod,
                                 std::vector<double>& eff) const {
    if (is_tip(v)) return;
    auto it = clades.find(v);
    if (it != clades.end()) clade_no = it->second;
    const double rate = rate_mod.get_rate_clade(clade_no);
    for (int c : {nodes_[v].left, nodes_[v].right}) {
        if (c < 0) continue;
        const double b = nodes_[c].branchlength;
        double len;
        if (v == root_ && !root_compat_ && rate_mod.get_n_cut_offs() == 0) len = b;
        else len = rate_mod.get_time_mod_branch_length(b, (float)dist);
        eff[c] = len * rate;
        effective_lengths(c, clade_no, dist + b, rate_mod, eff);
    }
}

bool simpleML::prepare(sub_model& model, rate_model& rate_mod, std::vector<int>& lib) {
    if (!ctx_) {
        const int rc = phylo_create(0, &ctx_);
        if (rc) { err_ = "no GPU available for libphylo_hip"; std::cerr << err_ << std::endl; return false; }
    }
    if (tipdata_.size() != nodes_.size()) tipdata_.resize(nodes_.size());
    const unsigned int S = model.get_n_states();
    const unsigned int m = get_n_char();
    const std::vector<int> tip = tips();
    if (m == 0 || tip.size() < 2) { err_ = "no characters set"; return false; }
    // check_nodes (simpleML.cpp:52-71): every tip needs a state at every character
    for (int v : tip) {
        const auto& row = tipdata_[v];
        for (unsigned int pos = 0; pos < m; ++pos) {
            bool got = false;
            if (pos < row.size())
                for (unsigned int i = 0; i < S; ++i) got = got || row[pos][i];
            if (!got) {
                std::cerr << nodes_[v].label << " lack character state for character" << pos << "!" << std::endl;
                err_ = "missing character state";
                return false;
            }
        }
    }
    // library numbering: tips first
    lib.assign(nodes_.size(), -1);
    int next = 0;
    for (int v : tip) lib[v] = next++;
    for (size_t v = 0; v < nodes_.size(); ++v)
        if (!is_tip((int)v)) lib[v] = next++;
    const int n_lib = next;
    std::vector<uint32_t> masks((size_t)tip.size() * m);
    for (size_t t = 0; t < tip.size(); ++t)
        for (unsigned int s = 0; s < m; ++s) masks[t * m + s] = (uint32_t)(tipdata_[tip[t]][s].to_ulong() & 0xFFFFFFFFul);
    // site patterns: the distinct columns, in order of first appearance
    std::vector<double> weights;
    unsigned int m_eval = m;
    site_pattern_.clear();
    if (patterns_) {
        std::map<std::vector<uint32_t>, unsigned int> index;
        std::vector<uint32_t> col(tip.size());
        site_pattern_.resize(m);
        for (unsigned int s = 0; s < m; ++s) {
            for (size_t t = 0; t < tip.size(); ++t) col[t] = masks[t * m + s];
            auto it = index.emplace(col, (unsigned int)index.size()).first;
            site_pattern_[s] = it->second;
            if (it->second == weights.size()) weights.push_back(0.0);
            weights[it->second] += 1.0;
        }
        m_eval = (unsigned int)weights.size();
        std::vector<uint32_t> packed((size_t)tip.size() * m_eval);
        for (unsigned int s = 0; s < m; ++s)
            for (size_t t = 0; t < tip.size(); ++t) packed[t * m_eval + site_pattern_[s]] = masks[t * m + s];
        masks.swap(packed);
    }
    n_patterns_ = m_eval;
    std::vector<int32_t> left(n_lib, -1), right(n_lib, -1);
    for (size_t v = 0; v < nodes_.size(); ++v) {
        if (lib[v] < 0 || is_tip((int)v)) continue;
        const int l = nodes_[v].left, r = nodes_[v].right;
        left[lib[v]] = l >= 0 ? lib[l] : (r >= 0 ? lib[r] : -1);
        right[lib[v]] = (l >= 0 && r >= 0) ? lib[r] : -1;
    }
    std::vector<double> eff_node(nodes_.size(), 0.0), eff(n_lib, 0.0);
    effective_lengths(root_, 0, 0.0, rate_mod, eff_node);
    for (size_t v = 0; v < nodes_.size(); ++v) eff[lib[v]] = eff_node[v];
    const std::vector<double> pi = model.frequencies();
    const std::vector<double> rw = model.root_weights();
    int builder = builder_;
    if (builder == P_AUTO) builder = model.is_tr() ? P_EIGEN : P_PADE;
    int rc = phylo_set_precision(ctx_, fp64_ ? PHYLO_PREC_FP64 : PHYLO_PREC_FP32);
    if (!rc) rc = phylo_set_tips(ctx_, (int)tip.size(), (int)m_eval, (int)S, masks.data());
    if (!rc && patterns_) rc = phylo_set_site_weights(ctx_, weights.data());
    if (!rc) rc = phylo_set_topology(ctx_, n_lib, lib[root_], left.data(), right.data());
    if (!rc) rc = phylo_set_edge_lengths(ctx_, eff.data());
    if (!rc) rc = phylo_set_model(ctx_, (int)S, model.Q().data(), rw.data(), builder == P_EIGEN ? pi.data() : nullptr,
                                  builder);
    if (!rc) rc = phylo_set_categories(ctx_, (int)cat_rates_.size(), cat_rates_.data(), cat_w_.data());
    if (rc) {
        err_ = phylo_last_error(ctx_);
        std::cerr << err_ << std::endl;
        return false;
    }
    n_states_ = S;
    return true;
}

double simpleML::calculate_log_likelihood(sub_model& model) {
    rate_model rate_mod;
    return calculate_log_likelihood(model, rate_mod);
}

double simpleML::calculate_log_likelihood(sub_model& model, rate_model& rate_mod) {
    if (!prepare(model, rate_mod, lib_of_node_)) return 0.0;
    std::vector<double> lnl(n_patterns_, 0.0);
    double total = 0.0;
    if (phylo_loglik(ctx_, lnl.data(), &total)) {
        err_ = phylo_last_error(ctx_);
        std::cerr << err_ << std::endl;
        return 0.0;
    }
    if (site_pattern_.empty()) {
        site_lnl_.swap(lnl);
    } else {
        site_lnl_.resize(site_pattern_.size());
        for (size_t s = 0; s < site_pattern_.size(); ++s) site_lnl_[s] = lnl[site_pattern_[s]];
    }
    return total;
}

std::vector<double> simpleML::nni_neighbourhood(sub_model& model, int mode, std::vector<int>* node,
                                                std::vector<int>* side) {
    rate_model rate_mod;
    std::vector<double> out;
    if (!prepare(model, rate_mod, lib_of_node_)) return out;
    int nv = 0;
    double base = 0.0;
    if (phylo_nni_score(ctx_, mode, 0, nullptr, nullptr, nullptr, &nv, &base)) return out;
    out.assign(nv, 0.0);
    std::vector<int32_t> nd(nv), sd(nv);
    if (phylo_nni_score(ctx_, mode, nv, out.data(), nd.data(), sd.data(), &nv, &base)) return {};
    std::vector<int> node_of_lib(nodes_.size(), -1);
    for (size_t v = 0; v < lib_of_node_.size(); ++v)
        if (lib_of_node_[v] >= 0) node_of_lib[lib_of_node_[v]] = (int)v;
    if (node) { node->clear(); for (int x : nd) node->push_back(x >= 0 ? node_of_lib[x] : -1); }
    if (side) side->assign(sd.begin(), sd.end());
    return out;
}

void simpleML::draw_normalized_likelihood_on_nodes() {
    if (!ctx_ || lib_of_node_.empty()) return;
    const unsigned int m = get_n_char(), S = n_states_;
    const int C = (int)cat_rates_.size();
    std::vector<double> clv((size_t)n_patterns_ * C * S);
    std::vector<int32_t> scale(n_patterns_);
    // the evaluated column of each site (the site itself without patterns)
    const auto col = [&](unsigned int s) { return site_pattern_.empty() ? s : site_pattern_[s]; };
    // postorder: children first, as the reference's recursion
    std::vector<int> order, stack{root_};
    std::vector<char> seen(nodes_.size(), 0);
    while (!stack.empty()) {
        const int v = stack.back();
        if (!seen[v]) {
            seen[v] = 1;
            if (nodes_[v].right >= 0) stack.push_back(nodes_[v].right);
            if (nodes_[v].left >= 0) stack.push_back(nodes_[v].left);
        } else {
            stack.pop_back();
            if (!is_tip(v)) order.push_back(v);
        }
    }
    for (int v : order) {
        if (phylo_node_partials(ctx_, lib_of_node_[v], clv.data(), scale.data())) continue;
        std::ostringstream lab;
        lab << "[&";
        double sum = 0.0;
        for (unsigned int s = 0; s < m; ++s) {
            if (s) lab << ",";
            lab << "P={";
            std::vector<double> x(S, 0.0);
            const unsigned int q = col(s);
            const double f = std::ldexp(1.0, scale[q]);
            for (unsigned int i = 0; i < S; ++i) {
                double a = 0.0;
                for (int c = 0; c < C; ++c) a += (C == 1 ? 1.0 : cat_w_[c]) * clv[((size_t)q * C + c) * S + i];
                x[i] = a * f;
                sum += x[i];
            }
            for (unsigned int i = 0; i < S; ++i) {
                if (i) lab << ',';
                lab << x[i] / sum;
            }
            lab << "}";
        }
        lab << "]";
        nodes_[v].label = (nodes_[v].has_label ? nodes_[v].label : std::string()) + lab.str();
        nodes_[v].has_label = true;
    }
}

// simulate_chars (simpleML.cpp:233-292): a root state drawn from charfreq,
// then on every edge a state drawn from row P[ancestor] of P(raw branch
// length), on the GPU (phylo_simulate).  Rows are returned for every labelled
// non-root node in preorder, as the reference does.
simulated_rows simpleML::simulate_chars(const std::vector<double>& charfreq, const unsigned int n_char,
                                        sub_model& model) {
    simulated_rows out;
    if (!ctx_ && phylo_create(0, &ctx_)) {
        err_ = "no GPU available for libphylo_hip";
        std::cerr << err_ << std::endl;
        return out;
    }
    const unsigned int S = model.get_n_states();
    const std::vector<int> tip = tips();
    if (n_char == 0 || tip.size() < 2) return out;
    std::vector<int> lib(nodes_.size(), -1);
    int next = 0;
    for (int v : tip) lib[v] = next++;
    for (size_t v = 0; v < nodes_.size(); ++v)
        if (!is_tip((int)v)) lib[v] = next++;
    const int n_lib = next;
    std::vector<int32_t> left(n_lib, -1), right(n_lib, -1);
    std::vector<double> len(n_lib, 0.0);
    for (size_t v = 0; v < nodes_.size(); ++v) {
        len[lib[v]] = nodes_[v].branchlength;
        if (is_tip((int)v)) continue;
        const int l = nodes_[v].left, r = nodes_[v].right;
        left[lib[v]] = l >= 0 ? lib[l] : (r >= 0 ? lib[r] : -1);
        right[lib[v]] = (l >= 0 && r >= 0) ? lib[r] : -1;
    }
    // the tip data is not used by the simulation; one all-state site fixes the numbering
    const std::vector<uint32_t> any(tip.size(), S >= 32 ? 0xFFFFFFFFu : ((1u << S) - 1u));
    const std::vector<double> pi = model.frequencies();
    const std::vector<double> rw = model.root_weights();
    int builder = builder_;
    if (builder == P_AUTO) builder = model.is_tr() ? P_EIGEN : P_PADE;
    std::vector<double> root_f(S, 0.0);
    for (unsigned int j = 0; j < S && j < charfreq.size(); ++j) root_f[j] = charfreq[j];
    std::vector<uint8_t> states((size_t)n_lib * n_char);
    int rc = phylo_set_tips(ctx_, (int)tip.size(), 1, (int)S, any.data());
    if (!rc) rc = phylo_set_topology(ctx_, n_lib, lib[root_], left.data(), right.data());
    if (!rc) rc = phylo_set_edge_lengths(ctx_, len.data());
    if (!rc) rc = phylo_set_model(ctx_, (int)S, model.Q().data(), rw.data(), builder == P_EIGEN ? pi.data() : nullptr,
                                  builder);
    if (!rc) rc = phylo_set_categories(ctx_, (int)cat_rates_.size(), cat_rates_.data(), cat_w_.data());
    if (!rc) rc = phylo_simulate(ctx_, (int)n_char, root_f.data(), seeded_ ? seed_ : std::random_device{}(),
                                 states.data());
    lib_of_node_.clear();       // the context no longer holds this tree's likelihood inputs
    if (rc) {
        err_ = phylo_last_error(ctx_);
        std::cerr << err_ << std::endl;
        return out;
    }
    std::vector<int> pre, stack{root_};
    while (!stack.empty()) {
        const int v = stack.back();
        stack.pop_back();
        pre.push_back(v);
        if (nodes_[v].right >= 0) stack.push_back(nodes_[v].right);
        if (nodes_[v].left >= 0) stack.push_back(nodes_[v].left);
    }
    for (int v : pre) {
        if (v == root_ || !nodes_[v].has_label || nodes_[v].label.empty()) continue;
        std::vector<state_set> row(n_char);
        const uint8_t* st = states.data() + (size_t)lib[v] * n_char;
        for (unsigned int c = 0; c < n_char; ++c) row[c].set(st[c]);
        out.taxa.push_back(nodes_[v].label);
        out.rows.push_back(std::move(row));
    }
    return out;
}

}  // namespace phylo_amd
