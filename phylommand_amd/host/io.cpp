// io.cpp -- see io.h.
#include "io.h"

#include <cctype>
#include <cstdlib>

namespace phylo_amd {

unsigned int character_vector::highest_char_state() const {
    unsigned int hi = 0;
    for (const auto& c : characters)
        for (unsigned int j = hi; j < SIZE; ++j)
            if (c.test(j)) hi = j;
    return hi;
}

void character_vector::set_empty_to(const state_set& value) {
    for (auto& c : characters)
        if (c.none()) c |= value;
}

namespace alphabet {

static void put(alphabet_map& a, char up, std::initializer_list<int> states) {
    state_set s;
    for (int k : states) s.set(k);
    a[up] = s;
    a[(char)std::tolower((unsigned char)up)] = s;
}

void set_alphabet_dna(alphabet_map& a) {
    put(a, 'A', {0}); put(a, 'G', {1}); put(a, 'C', {2}); put(a, 'T', {3});
    put(a, 'R', {0, 1}); put(a, 'Y', {2, 3}); put(a, 'S', {1, 2}); put(a, 'W', {0, 3});
    put(a, 'K', {1, 3}); put(a, 'M', {0, 2}); put(a, 'B', {1, 2, 3}); put(a, 'D', {0, 1, 3});
    put(a, 'H', {0, 2, 3}); put(a, 'V', {0, 1, 2}); put(a, 'N', {0, 1, 2, 3});
    a['-'] = state_set();
    a['.'] = state_set();
}

void set_alphabet_amino_acid(alphabet_map& a) {
    const char* order = "ABCDEFGHIKLMNPQRSTVWXYZ";
    for (int i = 0; order[i]; ++i) put(a, order[i], {i});
    a['-'] = state_set();
    a['.'] = state_set();
}

void set_alphabet_binary(alphabet_map& a) {
    state_set z, o, g;
    z.set(0);
    o.set(1);
    g.set(0);
    g.set(1);
    a['0'] = z;
    a['1'] = o;
    a['-'] = g;
}

void parse_alphabet_file(std::istream& in, alphabet_map& a) {
    char c, type = ' ';
    std::string value;
    state_set trait;
    char mode = 'b';
    while (in.get(c)) {
        if (mode == 'b' && (c == ' ' || c == '\n' || c == '\r' || c == '\t')) continue;
        if (mode == 'b') mode = 't';
        if (mode == 't') {
            if ((c == ' ' || c == '\t') && type != ' ') mode = 'c';
            else { type = c; mode = 'c'; }
        } else if (mode == 'c') {
            if ((c == ' ' || c == '\t' || c == '\n' || c == '\r') && type != ' ' && !value.empty()) {
                const unsigned int number = (unsigned int)atoi(value.c_str());
                if (number < SIZE) trait.set(number);
                value.clear();
            }
            if ((c == '\n' || c == '\r') && type != ' ') {
                a[type] = trait;
                trait.reset();
                value.clear();
                mode = 'b';
            } else if (c >= '0' && c <= '9') {
                value += c;
            } else if (c == '\n' || c == '\r') {
                mode = 'b';
            }
        }
    }
    if (mode == 'c' && type != ' ') {           // last line without a newline
        if (!value.empty()) {
            const unsigned int number = (unsigned int)atoi(value.c_str());
            if (number < SIZE) trait.set(number);
        }
        a[type] = trait;
    }
}

state_set all_states(const alphabet_map& a) {
    state_set s;
    for (const auto& kv : a) s |= kv.second;
    return s;
}

char translate_bitset(const state_set& s, const alphabet_map& a) {
    for (const auto& kv : a)
        if (kv.second == s) return kv.first;
    return '!';
}

}  // namespace alphabet

static state_set lookup(const alphabet_map& a, char c) {
    auto it = a.find(c);
    return it == a.end() ? state_set() : it->second;
}

void parse_fasta(std::istream& in, const alphabet_map& a, std::vector<character_vector>& out) {
    std::string text((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    character_vector row;
    std::string taxon;
    char mode = 'n';
    bool open = false;
    // the reference's read loop processes the last character once more when
    // the text does not end in a newline (matrix_parser.cpp:105-108)
    const size_t n = text.size();
    for (size_t i = 0; i <= n; ++i) {
        if (i == n && n == 0) break;
        const char c = i < n ? text[i] : text[n - 1];
        if (c == '>') {
            mode = 't';
        } else if (mode == 't' && (c == '\n' || c == '\r')) {
            if (open && !row.empty()) out.push_back(row);
            row.reset();
            row.set_taxon(taxon);
            open = true;
            taxon.clear();
            mode = 's';
        } else if (c != '\n' && c != '\r' && c != '\t' && c != ' ') {
            if (mode == 't') taxon += c;
            else if (mode == 's') row.add_character(lookup(a, c));
        }
    }
    if (open && !row.empty()) out.push_back(row);
}

void parse_relaxed_phylip(std::istream& in, const alphabet_map& a, std::vector<character_vector>& out) {
    unsigned int n_taxa = 0, n_char = 0;
    if (!(in >> n_taxa >> n_char)) return;
    for (unsigned int t = 0; t < n_taxa; ++t) {
        character_vector row;
        std::string name;
        if (!(in >> name)) break;
        row.set_taxon(name);
        char c;
        while (row.n_char() < n_char && in.get(c)) {
            if (c == ' ' || c == '\t' || c == '\n' || c == '\r') continue;
            row.add_character(lookup(a, c));
        }
        out.push_back(row);
    }
}

unsigned int n_states_of(const std::vector<character_vector>& rows) {
    unsigned int n = 0;
    for (const auto& r : rows)
        if (r.n_char() && r.highest_char_state() + 1 > n) n = r.highest_char_state() + 1;
    return n;
}

}  // namespace phylo_amd
