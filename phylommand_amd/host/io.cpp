// io.cpp -- see io.h.
#include "io.h"

#include <cctype>
#include <cstdlib>

#include "phylo_tree.h"

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

namespace {

bool line_end(char c) { return c == '\n' || c == '\r'; }

// The rows of a relaxed-phylip or NEXUS matrix: each row is a taxon name
// (its characters up to the first space), then its characters up to the end
// of the line.  Only characters of the alphabet count; spaces and unknown
// letters are skipped (unlike fasta, where an unknown letter is an empty
// state set).  A row is kept when its line ends, so a last row without a
// newline is dropped.  In a NEXUS matrix ';' ends the matrix.  Line breaks
// between rows are skipped; in phylip mode '\r' joins the taxon name and tabs
// separate nothing (matrix_parser.cpp:27-91, 140-225).
void read_matrix_rows(std::istream& in, const alphabet_map& a, bool nexus, std::vector<character_vector>& out) {
    character_vector row;
    std::string name;
    bool in_chars = false;
    char c;
    while (in.get(c)) {
        if (nexus && c == ';') return;
        if (!in_chars) {
            const bool skip = nexus ? (c == '\n' || c == '\r' || c == '\t') : c == '\n';
            if (skip) continue;
            if (c != ' ') {
                name += c;
            } else if (!name.empty()) {
                row.set_taxon(name);
                name.clear();
                in_chars = true;
            }
        } else if (line_end(c)) {
            out.push_back(row);
            row.reset();
            in_chars = false;
        } else {
            auto it = a.find(c);
            if (it != a.end()) row.add_character(it->second);
        }
    }
}

std::string lower(std::string w) {
    for (char& ch : w) ch = (char)std::tolower((unsigned char)ch);
    return w;
}

// Skips to the NEXUS block "begin <name>;" (words compared case-blind,
// file_parser::move_to_next_block, file_parser.cpp:62-100).
bool find_nexus_block(std::istream& in, const std::string& name) {
    std::string w;
    while (in >> w)
        if (lower(w) == "begin" && in >> w && lower(w) == name + ";") return true;
    return false;
}

}  // namespace

void parse_relaxed_phylip(std::istream& in, const alphabet_map& a, std::vector<character_vector>& out) {
    // header "n_taxa n_char": digits until a space after the first digit,
    // then digits until the end of the line
    unsigned int n_taxa = 0, n_char = 0;
    int field = 0;
    char c;
    while (field < 2 && in.get(c)) {
        if (field == 0) {
            if (c == ' ' && n_taxa > 0) field = 1;
            else if (c != ' ') n_taxa = n_taxa * 10 + (unsigned int)(c - '0');
        } else if (line_end(c)) {
            field = 2;
        } else if (c != ' ') {
            n_char = n_char * 10 + (unsigned int)(c - '0');
        }
    }
    (void)n_taxa;
    (void)n_char;
    read_matrix_rows(in, a, false, out);
}

bool parse_nexus_data(std::istream& in, const alphabet_map& a, std::vector<character_vector>& out) {
    if (!find_nexus_block(in, "data")) return false;
    // commands up to MATRIX: words end at ' ', ';', '=' or a line break
    // (not at a tab); NTAX and NCHAR values are read but only reported on
    std::string w;
    char c;
    while (in.get(c)) {
        if (c == ' ' || c == ';' || c == '=' || line_end(c)) {
            if (!w.empty() && lower(w) == "matrix") break;
            w.clear();
        } else {
            w += c;
        }
    }
    read_matrix_rows(in, a, true, out);
    return true;
}

nexus_tree_reader::nexus_tree_reader(std::istream& in) : in_(in) {
    ok_ = find_nexus_block(in_, "trees");
    if (ok_) pending_ = next_command();
    if (pending_ == "translate") {
        read_translate();
        pending_ = next_command();
    }
}

// file_parser::read_next_nexus_command (file_parser.cpp:102-130): the next
// word that names a command, lower case.
std::string nexus_tree_reader::next_command() {
    std::string w;
    while (in_ >> w) {
        w = lower(w);
        if (w == "translate" || w == "tree" || w == "end;") return w;
    }
    return std::string();
}

// file_parser::read_translate_parameters (file_parser.cpp:216-274): "key
// value" pairs separated by ',' up to ';'; [comments] skipped, quoted stretches
// taken as they are, spaces trimmed from the end of a value.
void nexus_tree_reader::read_translate() {
    std::string key, value;
    char state = 'k';             // 'k' before a key, 'K' in it, 'v' in the value
    const auto blank = [](char ch) { return ch == ' ' || ch == '\n' || ch == '\t' || ch == '\r'; };
    const auto store = [&]() {
        while (!value.empty() && value.back() == ' ') value.pop_back();
        translation_[key] = value;
        key.clear();
        value.clear();
    };
    int ch = in_.get();
    while (in_) {
        const char t = (char)ch;
        if (t == '[') {
            int depth = 1;
            while (in_ && depth) {
                ch = in_.get();
                if (ch == ']') --depth;
                else if (ch == '[') ++depth;
            }
        } else if (state != 'v') {
            if (!blank(t)) { key += t; state = 'K'; }
            else if (state == 'K') state = 'v';
        } else if (t == ',') {
            store();
            state = 'k';
        } else if (t == '\'' || t == '"') {
            // a quoted stretch is taken as it is; the closing quote is then
            // examined again and so opens the next stretch, as the
            // reference's loop does (file_parser.cpp:253-261)
            ch = in_.get();
            while (in_ && (char)ch != t) {
                value += (char)ch;
                ch = in_.get();
            }
            continue;
        } else if (!(t == ' ' && value.empty()) && t != '\n' && t != '\t' && t != '\r') {
            value += t;
        }
        ch = in_.get();
        if (ch == ';') {
            store();
            break;
        }
    }
}

bool nexus_tree_reader::next(tree& t) {
    if (!ok_) return false;
    if (first_) first_ = false;
    else pending_ = next_command();
    if (pending_ != "tree") return false;
    char c;
    while (in_.get(c) && c != '=') {}       // file_parser::move_to_start_of_tree (file_parser.cpp:276-282)
    if (!in_) return false;
    t.tree_file_parser(in_, translation_, false);
    return !t.empty();
}

unsigned int n_states_of(const std::vector<character_vector>& rows) {
    unsigned int n = 0;
    for (const auto& r : rows)
        if (r.n_char() && r.highest_char_state() + 1 > n) n = r.highest_char_state() + 1;
    return n;
}

}  // namespace phylo_amd
