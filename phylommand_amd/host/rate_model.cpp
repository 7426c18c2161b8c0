// rate_model.cpp -- see rate_model.h.
#include "rate_model.h"

#include <cstdlib>
#include <string>
#include <vector>

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

// ---- argument grammar -----------------------------------------------------
//
// Both arguments are lists of tokens.  A token ends at ':' (it is then a rate
// and opens a new rate class), at the list separator (';' between taxon sets,
// ',' between cut-off times) or at the end of the argument.  At the end the
// final character stays in the last token unless it is ':', so a trailing
// separator becomes part of the last item ("2:a,b;" names the taxon set
// "a,b;"), which is how treeator's -V and time arguments are read
// (rate_model.cpp:129-176).  Numbers are read with atof and stored as float.
//
//   clade argument:  (RATE ':' | TAXA (';' | end))*   TAXA: a taxon set, kept as text
//   time argument:   (RATE ':' | TIME (',' | end))*
//
// Items after a rate belong to that rate's class; items before the first
// rate belong to class 0, which a taxon set creates with rate 1 if no clade
// rate exists yet, while a cut-off time there is an error.

namespace {

struct arg_token {
    std::string text;
    char end;        // ':', the list separator, or 0 at the end of the argument
};

std::vector<arg_token> split_rate_arg(const char* argument, char separator) {
    std::vector<arg_token> tokens;
    const std::string arg(argument ? argument : "");
    size_t start = 0;
    for (size_t i = 0; i < arg.size(); ++i) {
        const char ch = arg[i];
        if (ch == ':') {
            tokens.push_back({arg.substr(start, i - start), ':'});
            start = i + 1;
        } else if (i + 1 == arg.size()) {
            tokens.push_back({arg.substr(start), 0});
            start = arg.size();
        } else if (ch == separator) {
            tokens.push_back({arg.substr(start, i - start), separator});
            start = i + 1;
        }
    }
    return tokens;
}

float arg_number(const std::string& text) { return (float)std::atof(text.c_str()); }

}  // namespace

bool rate_model::pars_clade_rates_arg(const char* argument, std::vector<std::string>& taxon_sets) {
    unsigned int current = 0;                       // rate class of the taxon sets that follow
    for (const arg_token& tok : split_rate_arg(argument, ';')) {
        if (tok.end == ':') {
            current = get_n_clade_rates();
            set_rate(current, arg_number(tok.text));
            continue;
        }
        taxon_sets.push_back(tok.text);
        if (current >= rates.size()) set_rate(current, 1.0f);    // a taxon set before any rate
        if (!set_clade_rate((unsigned int)taxon_sets.size() - 1, current)) return false;
    }
    return get_n_clade_rates() > 0;
}

bool rate_model::pars_rates_and_times_arg(const char* argument) {
    unsigned int current = 0;                       // rate class of the cut-offs that follow
    for (const arg_token& tok : split_rate_arg(argument, ',')) {
        if (tok.end == ':') {
            current = get_n_time_rates();
            set_time_rate(current, arg_number(tok.text));
        } else if (!add_rate_for_time(current, arg_number(tok.text))) {
            return false;                           // a time before any rate
        }
    }
    return get_n_time_rates() > 0;
}

}  // namespace phylo_amd
