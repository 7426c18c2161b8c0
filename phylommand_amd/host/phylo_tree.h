// phylo_tree.h -- the parts of phylommand's tree class (tree.h:47-518) that the
// likelihood path uses: newick input with the reference's rules (comments,
// quoted labels, polytomies resolved with zero-length nodes, trees rooted on
// a leaf; tree.cpp:69-237), newick output with 6-decimal branch lengths
// (tree.cpp:239-280), taxon sets / most recent common ancestors (tree.h:288,
// tree.cpp:990-1047), shortest branch and longest root-to-tip distance, and
// nearest-neighbour interchange (tree.cpp:2468-2548).
#ifndef PHYLO_AMD_PHYLO_TREE_H
#define PHYLO_AMD_PHYLO_TREE_H

#include <iostream>
#include <map>
#include <string>
#include <vector>

namespace phylo_amd {

class tree {
  public:
    struct node {
        double branchlength = 0.0;
        std::string label;
        bool has_label = false;
        int left = -1, right = -1, parent = -1;
    };

    tree() { clear(); }
    virtual ~tree() = default;

    // Reads one tree from the stream (stops after its ';'), translating tip
    // labels through label_translation (a NEXUS Translate table).  With
    // read_extra_node_annotation a "[&...]" comment is kept and appended to
    // the node's label; other comments are skipped (tree.h:85-88,
    // tree.cpp:69-237).
    void tree_file_parser(std::istream& infile);
    void tree_file_parser(std::istream& infile, const std::map<std::string, std::string>& label_translation);
    void tree_file_parser(std::istream& infile, const std::map<std::string, std::string>& label_translation,
                          bool read_extra_node_annotation);
    void tree_file_parser(std::istream& infile, bool read_extra_node_annotation) {
        tree_file_parser(infile, {}, read_extra_node_annotation);
    }
    bool empty() const { return nodes_[root_].left < 0 && nodes_[root_].right < 0 && !nodes_[root_].has_label; }
    void print_newick(std::ostream& out = std::cout, bool include_br_length = true, bool int_node_label = true) const;
    void print_newick(bool include_br_length, bool int_node_label = true) const {
        print_newick(std::cout, include_br_length, int_node_label);
    }
    std::string newick(bool include_br_length = true, bool int_node_label = true) const;

    // NEXUS output (tree.h:121-137, tree.cpp:289-331): the TAXA block and the
    // head of the TREES block with its Translate table.  An empty
    // translate_taxa is filled with "1".."n" for the tips in left-first
    // order; a filled one is written as it is.
    void print_nexus_tree_intro(std::ostream& out, std::map<std::string, std::string>& translate_taxa) const;
    void print_nexus_tree_intro(std::map<std::string, std::string>& translate_taxa) const {
        print_nexus_tree_intro(std::cout, translate_taxa);
    }
    // "tree NAME = (...);" with tips written as their Translate keys (or
    // numbered from 1 when the table is empty), "[...]" annotations kept.
    void print_tree_to_nexus(std::ostream& out, const std::string& name, bool include_br_length, bool int_node_label,
                             const std::map<std::string, std::string>& translate_taxa) const;
    void print_tree_to_nexus(const std::string& name, bool include_br_length, bool int_node_label,
                             const std::map<std::string, std::string>& translate_taxa) const {
        print_tree_to_nexus(std::cout, name, include_br_length, int_node_label, translate_taxa);
    }
    void print_tree_to_nexus(const std::string& name, bool include_br_length, bool int_node_label) const {
        print_tree_to_nexus(std::cout, name, include_br_length, int_node_label, {});
    }
    void print_nexus_tree_end(std::ostream& out = std::cout) const { out << "End;" << std::endl; }
    unsigned int n_tips() const { return n_sub_tips(root_); }
    double shortest_branch() const;
    double longest_to_tip() const;
    // Each set is "taxon,taxon,..."; its MRCA gets the set's index
    // (tree.h:288-292, clade numbers as the reference assigns them).
    void add_taxon_sets(const std::vector<std::string>& taxon_sets);
    int most_recent_common_ancestor(const std::string& taxa) const;
    // tree::nni (tree.cpp:2520-2548): L and R are copies of this tree with the
    // left / right child of internal branch `branch_no` swapped.
    void nni(unsigned int branch_no, tree& L, tree& R) const;

    // node access
    int root() const { return root_; }
    const std::vector<node>& nodes() const { return nodes_; }
    bool is_tip(int v) const { return nodes_[v].left < 0 && nodes_[v].right < 0; }
    int find_taxon_tip(const std::string& taxon) const;
    std::vector<int> tips() const;                 // tips in left-first order
    const std::map<int, unsigned int>& clade_map() const { return clades; }

  protected:
    void clear();
    int new_node();
    unsigned int n_sub_tips(int v) const;
    int internal_branch_by_number(unsigned int branch_no) const;
    void interchange(int leaf, bool left);
    void write_subtree(std::string& out, int v, bool br, bool int_label) const;
    // newick of the subtree at v with tips written by tip_name(v, &counter)
    template <class TipName>
    void write_subtree_with(std::string& out, int v, bool br, bool int_label, TipName tip_name) const;

    std::vector<node> nodes_;
    int root_ = 0;
    std::map<int, unsigned int> clades;
};

}  // namespace phylo_amd

#endif
