// rate_model.h -- branch-length rate model with the interface of phylommand's
// rate_model (rate_model.h:29-84): piecewise rates along time from the root
// (cut-offs) and clade rate multipliers, stored as float like the reference
// (rate_model.h:78-80) so effective branch lengths round the same way.
#ifndef PHYLO_AMD_RATE_MODEL_H
#define PHYLO_AMD_RATE_MODEL_H

#include <string>
#include <vector>

namespace phylo_amd {

class rate_model {
  public:
    rate_model() = default;
    bool add_rate_for_time(const unsigned int rate_class, const float time);
    bool set_cut_off(const unsigned int cut_off, float value);
    void set_time_rate(const unsigned int rate_class, const float rate);
    bool set_clade_rate(const unsigned int clade, const unsigned int rate_class);
    void set_rate(const unsigned int rate_class, const float rate);
    bool set_parameter(const unsigned int n, const float rate);
    unsigned int get_n_rates() const { return (unsigned int)(rates_in_time.size() + rates.size()); }
    unsigned int get_n_clade_rates() const { return (unsigned int)rates.size(); }
    unsigned int get_n_time_rates() const { return (unsigned int)rates_in_time.size(); }
    unsigned int get_n_cut_offs() const { return (unsigned int)cut_offs.size(); }
    unsigned int get_n_clades() const { return (unsigned int)clade_rates.size(); }
    unsigned int get_n_parameters() const {
        return (unsigned int)(rates_in_time.size() + rates.size() + cut_offs.size());
    }
    float get_cut_off(const unsigned int n) const { return n < cut_offs.size() ? cut_offs[n].cut_off : 0.0f; }
    float get_rate_in_time(const unsigned int n) const {
        return n < cut_offs.size() ? rates_in_time[cut_offs[n].rate_class] : 1.0f;
    }
    float get_rate_clade(const unsigned int n) const {
        return n < clade_rates.size() ? rates[clade_rates[n]] : 1.0f;
    }
    // Integral of the piecewise rate over [d, d + b] (rate_model.cpp:103-127),
    // including the reference's result for d == 0 with no cut-offs (0: the
    // root's child edges, rate_model.cpp:106,122).
    double get_time_mod_branch_length(const double branch_length, const float distance_to_root) const;
    // "rate:taxon,taxon;rate:..." (rate_model.cpp:129-151)
    bool pars_clade_rates_arg(const char* argument, std::vector<std::string>& taxon_sets);
    // "rate:time,time,rate:time" (rate_model.cpp:153-176)
    bool pars_rates_and_times_arg(const char* argument);

  private:
    struct rate_cut_off {
        unsigned int rate_class;
        float cut_off;
    };
    std::vector<rate_cut_off> cut_offs;   // sorted by time
    std::vector<float> rates_in_time;
    std::vector<float> rates;
    std::vector<unsigned int> clade_rates;
};

}  // namespace phylo_amd

#ifdef PHYLO_AMD_AS_REFERENCE
// Building phylommand's own sources against these classes (the treeator
// binding): the unqualified name they use.
using phylo_amd::rate_model;
#endif

#endif
