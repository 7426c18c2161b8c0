// sub_model.cpp -- see sub_model.h.
#include "sub_model.h"

#include <iostream>

#include "../csrc/model_math.h"

namespace phylo_amd {

sub_model::sub_model(unsigned int dim, bool tr, bool eq)
    : timerev(tr), equal_freq(eq), n_states(dim), rate_specs(n_posible_rates(), 0), rates(1, 0.0),
      state_freqs(dim, 0.0), Qm((size_t)dim * dim, 0.0) {
    calc_Q_matrix();
}

sub_model::sub_model(unsigned int dim, bool tr) : sub_model(dim, tr, false) {}

// get_rate_pos (sub_model.cpp:3-29): n_posible_rates() for the diagonal.
unsigned int sub_model::rate_pos(const unsigned int row, const unsigned int col) const {
    if (row == col || row >= n_states || col >= n_states) return n_posible_rates();
    if (timerev) {
        const unsigned int i = row < col ? row : col, d = row < col ? col - row : row - col;
        // rows 0..i-1 of the upper triangle hold (n-1) + (n-2) + ... + (n-i) slots
        unsigned int pos = 0;
        for (unsigned int k = 1; k <= i; ++k) pos += n_states - k;
        return pos + d - 1;
    }
    return row * (n_states - 1) + col - (col > row ? 1 : 0);
}

double sub_model::get_freq(const unsigned int n) const {
    if (n >= n_states) return 0.0;
    if (equal_freq || !freqs_given) return 1.0 / n_states;
    if (n == n_states - 1) {
        double f = 1.0;
        for (unsigned int i = 0; i + 1 < n_states; ++i) f -= state_freqs[i];
        return f;
    }
    return state_freqs[n];
}

std::vector<double> sub_model::frequencies() const {
    std::vector<double> f(n_states);
    for (unsigned int i = 0; i < n_states; ++i) f[i] = get_freq(i);
    return f;
}

std::vector<double> sub_model::root_weights() const {
    if (timerev) return frequencies();
    return std::vector<double>(n_states, 1.0);
}

bool sub_model::set_freq(const unsigned int n, const double value) {
    if (n >= n_states || value < 0.0) return false;
    double sum = 0.0;
    for (unsigned int i = 0; i + 1 < n_states; ++i) sum += (i == n) ? value : state_freqs[i];
    if (sum > 1.0) return false;
    state_freqs[n] = value;
    freqs_given = true;
    calc_Q_matrix();
    return true;
}

void sub_model::set_freq_same() {
    for (unsigned int i = 0; i < n_states; ++i) state_freqs[i] = 1.0 / n_states;
    freqs_given = true;
    calc_Q_matrix();
}

// set_rate_spec (sub_model.cpp:55-97): parameters must be numbered in order,
// i.e. every parameter below `param` is already used by another position.
bool sub_model::set_rate_spec(const unsigned int n, const unsigned int param) {
    const unsigned int npos = n_posible_rates();
    if (n >= npos) return false;
    std::vector<char> used(npos + 1, 0);
    unsigned int mx = 0;
    for (unsigned int i = 0; i < npos; ++i) {
        if (i == n) continue;
        if (rate_specs[i] <= npos) used[rate_specs[i]] = 1;
        if (rate_specs[i] > mx) mx = rate_specs[i];
    }
    for (unsigned int i = 0; i < param && i <= npos; ++i)
        if (!used[i]) return false;
    if (param > mx) mx = param;
    rate_specs[n] = param;
    if (mx + 1 > rates.size()) rates.resize(mx + 1, 0.0);
    return true;
}

bool sub_model::set_rate_spec(const unsigned int row, const unsigned int col, const unsigned int param) {
    return set_rate_spec(rate_pos(row, col), param);
}

bool sub_model::set_rate(const unsigned int n, const double value) {
    if (n >= rates.size() || !(value >= 0.0)) return false;
    rates[n] = value;
    calc_Q_matrix();
    return true;
}

bool sub_model::set_rate(const unsigned int row, const unsigned int col, const double value) {
    const unsigned int pos = rate_pos(row, col);
    if (pos >= n_posible_rates()) return false;
    return set_rate(rate_specs[pos], value);
}

double sub_model::get_rate(const unsigned int n) const { return n < rates.size() ? rates[n] : 0.0; }

double sub_model::get_rate(const unsigned int row, const unsigned int col) const {
    const unsigned int pos = rate_pos(row, col);
    if (pos >= n_posible_rates()) return 0.0;
    return rates[rate_specs[pos]];
}

void sub_model::calc_Q_matrix() {
    for (unsigned int i = 0; i < n_states; ++i) {
        double neg = 0.0;
        for (unsigned int j = 0; j < n_states; ++j) {
            if (i == j) continue;
            double v = get_rate(i, j);
            if (timerev) v *= get_freq(j);
            neg -= v;
            Qm[(size_t)i * n_states + j] = v;
        }
        Qm[(size_t)i * n_states + i] = neg;
    }
}

void sub_model::set_P_matrix(const double branch_length) {
    Pm.assign((size_t)n_states * n_states, 0.0);
    phylo_math::p_refcompat((int)n_states, Qm.data(), branch_length, Pm.data());
}

double sub_model::get_P_value(const unsigned int row, const unsigned int col) const {
    if (row < n_states && col < n_states && !Pm.empty()) return Pm[(size_t)row * n_states + col];
    return 0.0;
}

void sub_model::print_Q_matrix(std::ostream& out) const {
    for (unsigned int x = 0; x < n_states; ++x) {
        for (unsigned int y = 0; y < n_states; ++y) {
            out << Qm[(size_t)x * n_states + y];
            if (y + 1 < n_states) out << ' ';
        }
        if (x + 1 < n_states) out << std::endl;
    }
}

void sub_model::print_Q_matrix() const { print_Q_matrix(std::cout); }

}  // namespace phylo_amd
