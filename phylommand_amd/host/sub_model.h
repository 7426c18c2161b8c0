// sub_model.h -- substitution model with the interface of phylommand's
// sub_model (sub_model.h:28-128): rate specifications over the reference's
// rate positions, rates, state frequencies, Q, and P(t).
//
// Semantics kept from the reference:
//  - Q_ij = rate[spec[pos(i,j)]] * (reversible ? pi_j : 1), Q_ii = -row sum,
//    unnormalised (sub_model.cpp:99-114); pos() walks the upper triangle
//    row-major for reversible models and all off-diagonals otherwise
//    (sub_model.cpp:3-29).
//  - The last frequency is 1 - sum of the others (sub_model.cpp:44-53);
//    set_freq rejects values that make the others sum past 1 (:31-42).
//  - set_P_matrix(t) builds P = (I + Q t / 2^20)^(2^20) like the reference
//    (marth.cpp:129-140), so get_P_value matches it bit for bit; the
//    likelihood engine uses exact P(t) unless asked for this builder.
// Memory-safety defects of the reference are not reproduced: any number of
// distinct rates is allowed (the reference overflows its heap from four,
// sub_model.cpp:87).  Frequencies not yet set count as 0 in set_freq's sum
// check, as the reference's freshly allocated array does, and until any
// frequency is set the model uses equal frequencies (the reference reads
// uninitialised memory there, sub_model.h:85-88 with set_rate).
#ifndef PHYLO_AMD_SUB_MODEL_H
#define PHYLO_AMD_SUB_MODEL_H

#include <ostream>
#include <vector>

namespace phylo_amd {

class sub_model {
  public:
    sub_model(unsigned int dim, bool tr, bool eq);
    explicit sub_model(unsigned int dim, bool tr = false);

    void set_P_matrix(const double branch_length);
    double get_P_value(const unsigned int row, const unsigned int col) const;
    bool set_rate(const unsigned int n, const double value);
    bool set_rate(const unsigned int row, const unsigned int col, const double value);
    double get_rate(const unsigned int n) const;
    double get_rate(const unsigned int row, const unsigned int col) const;
    bool set_rate_spec(const unsigned int n, const unsigned int param);
    bool set_rate_spec(const unsigned int row, const unsigned int col, const unsigned int param);
    unsigned int get_n_rates() const { return (unsigned int)rates.size(); }
    bool set_freq(const unsigned int n, const double value);
    void set_freq_same();
    double get_freq(const unsigned int n) const;
    void equal_state_freq(const bool eq) { equal_freq = eq; calc_Q_matrix(); }
    bool is_time_reversible() const { return timerev; }
    unsigned int get_n_states() const { return n_states; }
    bool is_timerev() const { return timerev; }
    bool is_tr() const { return timerev; }
    bool are_state_freq_eqal() const { return equal_freq; }
    unsigned int n_posible_rates() const {
        return timerev ? (n_states * n_states - n_states) / 2 : n_states * n_states - n_states;
    }
    void print_Q_matrix(std::ostream& out) const;
    void print_Q_matrix() const;

    // Additions used by the GPU engine.
    const std::vector<double>& Q() const { return Qm; }
    std::vector<double> frequencies() const;
    // Weights of the root sum: pi for reversible models, ones otherwise
    // (simpleML.h:132-133).
    std::vector<double> root_weights() const;
    unsigned int rate_pos(const unsigned int row, const unsigned int col) const;

  private:
    void calc_Q_matrix();
    bool timerev;
    bool equal_freq;
    bool freqs_given = false;
    unsigned int n_states;
    std::vector<unsigned int> rate_specs;
    std::vector<double> rates;
    std::vector<double> state_freqs;
    std::vector<double> Qm;
    std::vector<double> Pm;
};

}  // namespace phylo_amd

#ifdef PHYLO_AMD_AS_REFERENCE
// Building phylommand's own sources against these classes (the treeator
// binding): the unqualified name they use.
using phylo_amd::sub_model;
#endif

#endif
