// rate_model.cpp -- see rate_model.h.
#include "rate_model.h"

#include <cstdlib>

namespace phylo_amd {

bool rate_model::add_rate_for_time(const unsigned int rate_class, const float time) {
    if (rate_class >= rates_in_time.size()) return false;
    // insert before the first later cut-off (equal times keep insertion order)
    size_t pos = 0;
    while (pos < cut_offs.size() && !(cut_offs[pos].cut_off > time)) ++pos;
    cut_offs.insert(cut_offs.begin() + pos, rate_cut_off{rate_class, time});
    return true;
}

bool rate_model::set_cut_off(const unsigned int cut_off, float value) {
    if (cut_off >= cut_offs.size()) return false;
    if (cut_off > 0 && value < cut_offs[cut_off - 1].cut_off) return false;
    if (cut_off + 1 < cut_offs.size() && value > cut_offs[cut_off + 1].cut_off) return false;
    cut_offs[cut_off].cut_off = value;
    return true;
}

void rate_model::set_time_rate(const unsigned int rate_class, const float rate) {
    if (rate_class >= rates_in_time.size()) rates_in_time.resize(rate_class + 1, rate);
    else rates_in_time[rate_class] = rate;
}

bool rate_model::set_clade_rate(const unsigned int clade, const unsigned int rate_class) {
    if (rate_class >= rates.size()) return false;
    if (clade >= clade_rates.size()) clade_rates.resize(clade + 1, rate_class);
    else clade_rates[clade] = rate_class;
    return true;
}

void rate_model::set_rate(const unsigned int rate_class, const float rate) {
    if (rate_class >= rates.size()) rates.resize(rate_class + 1, rate);
    else rates[rate_class] = rate;
}

bool rate_model::set_parameter(const unsigned int n, const float rate) {
    const unsigned int nc = (unsigned int)cut_offs.size(), nt = (unsigned int)rates_in_time.size();
    if (n < nc) return set_cut_off(n, rate);
    if (n - nc < nt) { set_time_rate(n - nc, rate); return true; }
    if (n - nc - nt < rates.size()) { set_rate(n - nc - nt, rate); return true; }
    return false;
}

double rate_model::get_time_mod_branch_length(const double branch_length, const float distance_to_root) const {
    const double d = distance_to_root;
    const double end = d + branch_length;
    double len = 0.0;
    double prev = -0.0;
    for (const auto& c : cut_offs) {
        const float cut = c.cut_off;
        const float r = rates_in_time[c.rate_class];
        if (distance_to_root < cut && end > prev) {
            if (d >= prev && end <= cut) len += (end - d) * r;               // inside this interval
            else if (d < prev && end <= cut) len += (end - prev) * r;        // enters from the previous one
            else if (d >= prev && end > cut) len += (cut - distance_to_root) * r;   // leaves into the next one (float arithmetic, as the reference)
            else if (d < prev && end > cut) len += (cut - prev) * r;         // spans the whole interval
        }
        prev = cut;
    }
    if (distance_to_root < prev && end > prev) len += end - prev;            // beyond the last cut-off
    else if (distance_to_root > prev) len = branch_length;
    return len;
}

bool rate_model::pars_clade_rates_arg(const char* argument, std::vector<std::string>& taxon_sets) {
    unsigned int rate_class = 0;
    std::string temp;
    for (unsigned int j = 0; argument[j] != '\0'; ++j) {
        if (argument[j] == ':') {
            rate_class = get_n_clade_rates();
            set_rate(rate_class, (float)atof(temp.c_str()));
            temp.clear();
        } else if (argument[j] == ';' || argument[j + 1] == '\0') {
            if (argument[j + 1] == '\0') temp += argument[j];
            taxon_sets.push_back(temp);
            if (!set_clade_rate((unsigned int)taxon_sets.size() - 1, rate_class)) {
                set_rate(rate_class, 1.0f);
                if (!set_clade_rate((unsigned int)taxon_sets.size() - 1, rate_class)) return false;
            }
            temp.clear();
        } else {
            temp += argument[j];
        }
    }
    return get_n_clade_rates() > 0;
}

bool rate_model::pars_rates_and_times_arg(const char* argument) {
    std::string temp;
    unsigned int rate_class = 0;
    for (unsigned int i = 0; argument[i] != '\0'; ++i) {
        if (argument[i] == ':') {
            rate_class = get_n_time_rates();
            set_time_rate(rate_class, (float)atof(temp.c_str()));
            temp.clear();
        } else if (argument[i] == ',' || argument[i + 1] == '\0') {
            if (argument[i + 1] == '\0') temp += argument[i];
            if (!add_rate_for_time(rate_class, (float)atof(temp.c_str()))) return false;
            temp.clear();
        } else {
            temp += argument[i];
        }
    }
    return get_n_time_rates() > 0;
}

}  // namespace phylo_amd
