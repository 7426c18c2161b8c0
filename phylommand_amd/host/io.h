// io.h -- alignment ingestion for the host classes: character vectors,
// the built-in alphabets and whitespace alphabet files, and the fasta and
// relaxed-phylip readers, with phylommand's semantics (matrix_parser.cpp,
// character_vector.h).
#ifndef PHYLO_AMD_IO_H
#define PHYLO_AMD_IO_H

#include <bitset>
#include <istream>
#include <map>
#include <string>
#include <vector>

namespace phylo_amd {

static const unsigned int SIZE = 24;   // states per character (constants.h:25)
typedef std::bitset<SIZE> state_set;
typedef std::map<char, state_set> alphabet_map;

// A taxon's row of characters (character_vector.h:32-91).
class character_vector {
  public:
    void add_character(const state_set& c) { characters.push_back(c); }
    state_set& get_character(unsigned int pos) {
        if (pos >= characters.size()) characters.resize(pos + 1);
        return characters[pos];
    }
    const state_set& character(unsigned int pos) const { return characters[pos]; }
    unsigned int n_char() const { return (unsigned int)characters.size(); }
    const std::string& get_taxon() const { return taxon; }
    void set_taxon(const std::string& name) { taxon = name; }
    bool empty() const { return characters.empty() && taxon.empty(); }
    void reset() { characters.clear(); taxon.clear(); }
    // highest set state over all characters (character_vector.h:75-82)
    unsigned int highest_char_state() const;
    // empty characters (gaps, unknown letters) become `value` (:83-86)
    void set_empty_to(const state_set& value);

  private:
    std::vector<state_set> characters;
    std::string taxon;
};

namespace alphabet {
void set_alphabet_dna(alphabet_map& a);          // A=0 G=1 C=2 T=3 + IUPAC (matrix_parser.cpp:282-323)
void set_alphabet_amino_acid(alphabet_map& a);   // 23 letters (matrix_parser.cpp:324-375)
void set_alphabet_binary(alphabet_map& a);       // 0, 1, - (matrix_parser.cpp:275-281)
// "X n n n" lines (alphabet_parser::pars_whitespace, matrix_parser.cpp:227-272)
void parse_alphabet_file(std::istream& in, alphabet_map& a);
state_set all_states(const alphabet_map& a);
char translate_bitset(const state_set& s, const alphabet_map& a);
}  // namespace alphabet

// matrix_parser::pars_fasta (matrix_parser.cpp:93-139).  Unknown letters are
// empty state sets; call set_empty_to to turn them into "any state".
void parse_fasta(std::istream& in, const alphabet_map& a, std::vector<character_vector>& out);
// matrix_parser::pars_relaxed_phylip (matrix_parser.cpp:27-91): a
// "n_taxa n_char" header, then "name sequence" rows, one per line; only
// letters of the alphabet are kept and a row is kept when its line ends.
void parse_relaxed_phylip(std::istream& in, const alphabet_map& a, std::vector<character_vector>& out);
// A NEXUS DATA block: file_parser::move_to_next_X_block(DATA) then
// matrix_parser::pars_nexus (file_parser.cpp:62-100, matrix_parser.cpp:140-225):
// commands up to MATRIX, then rows as in relaxed phylip up to ';'.
// Returns false when the stream has no DATA block.
bool parse_nexus_data(std::istream& in, const alphabet_map& a, std::vector<character_vector>& out);

class tree;
// The trees of a NEXUS TREES block, read as treeator does (treeator.cpp:490-530
// with file_parser.cpp:102-283): an optional Translate table first, then one
// tree per TREE command, tip labels translated through the table.
class nexus_tree_reader {
  public:
    explicit nexus_tree_reader(std::istream& in);
    bool found() const { return ok_; }                     // a TREES block exists
    const std::map<std::string, std::string>& translation() const { return translation_; }
    // the next tree into t; false after the last
    bool next(tree& t);

  private:
    std::string next_command();
    void read_translate();
    std::istream& in_;
    std::map<std::string, std::string> translation_;
    std::string pending_;
    bool ok_ = false, first_ = true;
};

// Number of states the data need (treeator.cpp:555-558).
unsigned int n_states_of(const std::vector<character_vector>& rows);

}  // namespace phylo_amd

#endif
