// tree.cpp -- see tree.h.
#include "phylo_tree.h"

#include <cstdio>
#include <cstdlib>
#include <limits>

namespace phylo_amd {

void tree::clear() {
    nodes_.assign(1, node());
    root_ = 0;
    clades.clear();
}

int tree::new_node() {
    nodes_.push_back(node());
    return (int)nodes_.size() - 1;
}

void tree::tree_file_parser(std::istream& infile) { tree_file_parser(infile, {}, false); }

void tree::tree_file_parser(std::istream& infile, const std::map<std::string, std::string>& tr) {
    tree_file_parser(infile, tr, false);
}

void tree::tree_file_parser(std::istream& infile, const std::map<std::string, std::string>& tr, bool keep_extra) {
    clear();
    int present = root_;
    std::string label, extra, blen;
    char mode = 's', c;
    auto flush = [&](int v) {
        if (!label.empty() || !extra.empty()) {
            auto it = tr.find(label);
            nodes_[v].label = ((it != tr.end()) ? it->second : label) + extra;
            nodes_[v].has_label = true;
            label.clear();
            extra.clear();
        }
        if (!blen.empty()) {
            nodes_[v].branchlength = atof(blen.c_str());
            blen.clear();
        }
    };
    bool ok = (bool)(infile >> c);
    while (ok) {
        if (c == '[') {                                  // comments, possibly nested
            ok = (bool)(infile >> c);
            if (!ok) break;
            if (c == '&' && keep_extra) {                // "[&...]": kept, joins the label
                extra += '[';
                while (ok && c != ']') {
                    extra += c;
                    ok = (bool)(infile >> c);
                }
                if (ok) extra += ']';
            } else if (c != ']') {
                int depth = 1;
                while (ok && depth > 0) {
                    if (c == '[') ++depth;
                    if (c == ']') --depth;
                    ok = (bool)(infile >> c);
                }
                continue;
            }
        } else if (c == '(') {
            mode = 'l';
            const int k = new_node();
            nodes_[present].left = k;
            nodes_[k].parent = present;
            present = k;
            label.clear();
            extra.clear();
        } else if (c == ',') {
            mode = 'l';
            flush(present);
            present = nodes_[present].parent;
            if (present < 0) { clear(); return; }
            if (nodes_[present].right < 0) {
                const int k = new_node();
                nodes_[present].right = k;
                nodes_[k].parent = present;
                present = k;
            } else {                                     // polytomy: zero-length node
                const int e = new_node();
                nodes_[e].left = nodes_[present].left;
                nodes_[e].right = nodes_[present].right;
                nodes_[nodes_[e].left].parent = e;
                nodes_[nodes_[e].right].parent = e;
                nodes_[e].parent = present;
                nodes_[present].left = e;
                const int k = new_node();
                nodes_[present].right = k;
                nodes_[k].parent = present;
                present = k;
            }
        } else if (c == ')') {
            mode = 'l';
            flush(present);
            present = nodes_[present].parent;
            if (present < 0) { clear(); return; }
        } else if (c == ':') {
            mode = 'b';
        } else if (c == ';') {
            if (present == root_) {
                flush(present);
                node& r = nodes_[root_];
                if (r.right < 0 && r.left >= 0 && r.has_label) {   // rooted on a leaf
                    const int k = new_node();
                    node& rr = nodes_[root_];
                    rr.right = k;
                    nodes_[k].parent = root_;
                    nodes_[k].label = rr.label;
                    nodes_[k].has_label = true;
                    nodes_[k].branchlength = rr.branchlength;
                    rr.label.clear();
                    rr.has_label = false;
                    rr.branchlength = 0.0;
                }
            } else {
                clear();                                 // unbalanced brackets
            }
            return;
        } else if (mode == 'b') {
            blen += c;
        } else if (mode == 'l') {
            if (label.empty() && (c == '\'' || c == '"')) {
                const char marker = c;
                ok = (bool)(infile >> c);
                while (ok && c != marker) {
                    label += c;
                    ok = (bool)(infile >> c);
                }
            } else {
                label += c;
            }
        }
        ok = (bool)(infile >> c);
    }
    if (present != root_) clear();
}

void tree::write_subtree(std::string& out, int v, bool br, bool int_label) const {
    const node& n = nodes_[v];
    if (n.left >= 0) {
        out += '(';
        write_subtree(out, n.left, br, int_label);
    } else if (n.has_label) {
        out += n.label;
    }
    if (n.right >= 0) {
        out += ',';
        write_subtree(out, n.right, br, int_label);
        out += ')';
        if (int_label && n.has_label) out += n.label;
    }
    if (v != root_ && br) {
        char buf[64];
        snprintf(buf, sizeof buf, ":%f", n.branchlength);
        out += buf;
    }
}

template <class TipName>
void tree::write_subtree_with(std::string& out, int v, bool br, bool int_label, TipName tip_name) const {
    const node& n = nodes_[v];
    if (n.left >= 0) {
        out += '(';
        write_subtree_with(out, n.left, br, int_label, tip_name);
    } else {
        out += tip_name(v);
    }
    if (n.right >= 0) {
        out += ',';
        write_subtree_with(out, n.right, br, int_label, tip_name);
        out += ')';
        if (int_label && n.has_label) out += n.label;
    }
    if (v != root_ && br) {
        char buf[64];
        snprintf(buf, sizeof buf, ":%f", n.branchlength);
        out += buf;
    }
}

namespace {
// A label without its "[...]" annotation, and the annotation alone.
std::string bare_name(const std::string& label) { return label.substr(0, label.find('[')); }
std::string annotation(const std::string& label) {
    const size_t k = label.find('[');
    return k == std::string::npos ? std::string() : label.substr(k);
}
}  // namespace

void tree::print_nexus_tree_intro(std::ostream& out, std::map<std::string, std::string>& translate_taxa) const {
    out << "#NEXUS" << std::endl << "Begin taxa;" << std::endl;
    out << "\tDimensions ntax=" << n_tips() << ";" << std::endl << "\tTaxlabels" << std::endl;
    std::vector<std::pair<std::string, std::string>> rows;       // (key, name) in output order
    if (translate_taxa.empty()) {
        int k = 0;
        for (int v : tips()) {
            const std::string name = nodes_[v].has_label ? bare_name(nodes_[v].label) : std::string();
            rows.emplace_back(std::to_string(++k), name);
            translate_taxa[rows.back().first] = name;
        }
    } else {
        for (const auto& kv : translate_taxa) rows.emplace_back(kv.first, kv.second);
    }
    for (const auto& r : rows) out << "\t\t" << r.second << std::endl;
    out << ";" << std::endl << "End;" << std::endl << std::endl;
    out << "Begin trees;" << std::endl << "\tTranslate" << std::endl;
    for (size_t i = 0; i < rows.size(); ++i) {
        out << "\t\t" << rows[i].first << ' ' << rows[i].second;
        if (i + 1 < rows.size()) out << ',' << std::endl;
    }
    out << std::endl << ";" << std::endl;
}

void tree::print_tree_to_nexus(std::ostream& out, const std::string& name, bool br, bool int_label,
                               const std::map<std::string, std::string>& translate_taxa) const {
    int counter = 1;
    std::string s;
    write_subtree_with(s, root_, br, int_label, [&](int v) {
        const node& n = nodes_[v];
        if (translate_taxa.empty())
            return std::to_string(counter++) + (n.has_label ? annotation(n.label) : std::string());
        if (!n.has_label) return std::string();
        const std::string bare = bare_name(n.label);
        std::string key;
        for (const auto& kv : translate_taxa)
            if (kv.second == bare) { key = kv.first; break; }
        return key + annotation(n.label);
    });
    out << "tree " << name << " = " << s << ";" << std::endl;
}

std::string tree::newick(bool br, bool int_label) const {
    std::string s;
    write_subtree(s, root_, br, int_label);
    return s + ";";
}

void tree::print_newick(std::ostream& out, bool br, bool int_label) const { out << newick(br, int_label) << std::endl; }

unsigned int tree::n_sub_tips(int v) const {
    if (v < 0) return 0;
    if (is_tip(v)) return 1;
    return n_sub_tips(nodes_[v].left) + n_sub_tips(nodes_[v].right);
}

std::vector<int> tree::tips() const {
    std::vector<int> out, stack{root_};
    while (!stack.empty()) {
        const int v = stack.back();
        stack.pop_back();
        if (is_tip(v)) { out.push_back(v); continue; }
        if (nodes_[v].right >= 0) stack.push_back(nodes_[v].right);
        if (nodes_[v].left >= 0) stack.push_back(nodes_[v].left);
    }
    return out;
}

int tree::find_taxon_tip(const std::string& taxon) const {
    for (int v : tips())
        if (nodes_[v].has_label && nodes_[v].label == taxon) return v;
    return -1;
}

double tree::shortest_branch() const {
    double best = std::numeric_limits<double>::infinity();
    for (size_t v = 0; v < nodes_.size(); ++v)
        if ((int)v != root_ && nodes_[v].branchlength < best) best = nodes_[v].branchlength;
    return best;
}

double tree::longest_to_tip() const {
    double best = 0.0;
    for (int v : tips()) {
        double d = 0.0;
        for (int u = v; u != root_ && u >= 0; u = nodes_[u].parent) d += nodes_[u].branchlength;
        if (d > best) best = d;
    }
    return best;
}

int tree::most_recent_common_ancestor(const std::string& taxa) const {
    std::vector<int> found;
    size_t start = 0;
    while (start <= taxa.size()) {
        size_t comma = taxa.find(',', start);
        if (comma == std::string::npos) comma = taxa.size();
        const int v = find_taxon_tip(taxa.substr(start, comma - start));
        if (v >= 0) found.push_back(v);
        start = comma + 1;
    }
    if (found.empty()) return -1;
    std::vector<int> path;
    for (int u = found[0]; u >= 0; u = nodes_[u].parent) path.push_back(u);
    size_t best = 0;   // index into path of the deepest common ancestor so far
    for (size_t k = 1; k < found.size(); ++k) {
        std::vector<char> on(nodes_.size(), 0);
        for (int u = found[k]; u >= 0; u = nodes_[u].parent) on[u] = 1;
        while (best < path.size() && !on[path[best]]) ++best;
    }
    return best < path.size() ? path[best] : root_;
}

void tree::add_taxon_sets(const std::vector<std::string>& taxon_sets) {
    for (const auto& s : taxon_sets) {
        const int v = most_recent_common_ancestor(s);
        if (v >= 0 && !clades.count(v)) {
            const unsigned int idx = (unsigned int)clades.size();
            clades[v] = idx;
        }
    }
}

// get_internal_branch_by_number_unrooted (tree.cpp:2455-2478): internal nodes
// numbered in left-first preorder, skipping the root when both of its
// subtrees hold two or more tips.
int tree::internal_branch_by_number(unsigned int branch_no) const {
    if (branch_no == 0) return -1;
    const int a = nodes_[root_].left, b = nodes_[root_].right;
    int start = root_;
    unsigned int number = branch_no + 1;
    if (n_sub_tips(a) < 2) { start = b; number = branch_no; }
    else if (n_sub_tips(b) < 2) { start = a; number = branch_no; }
    std::vector<int> stack;
    if (start >= 0) stack.push_back(start);
    unsigned int count = 0;
    while (!stack.empty()) {
        const int v = stack.back();
        stack.pop_back();
        if (v < 0 || is_tip(v)) continue;
        if (++count == number) return v;
        if (nodes_[v].right >= 0) stack.push_back(nodes_[v].right);
        stack.push_back(nodes_[v].left);
    }
    return -1;
}

void tree::interchange(int leaf, bool left) {
    if (leaf < 0 || nodes_[leaf].parent < 0 || is_tip(leaf)) return;
    const int par = nodes_[leaf].parent;
    int partner = -1;
    if (nodes_[par].left >= 0 && nodes_[par].left != leaf) {
        if (par != root_) partner = nodes_[par].left;
        else if (nodes_[nodes_[par].left].left >= 0) partner = nodes_[nodes_[par].left].left;
    } else if (nodes_[par].right >= 0 && nodes_[par].right != leaf) {
        if (par != root_) partner = nodes_[par].right;
        else if (nodes_[nodes_[par].right].left >= 0) partner = nodes_[nodes_[par].right].left;
    }
    if (partner < 0) return;
    int tmp = -1;
    if (left && nodes_[leaf].left >= 0) { tmp = nodes_[leaf].left; nodes_[leaf].left = partner; }
    if (!left && nodes_[leaf].right >= 0) { tmp = nodes_[leaf].right; nodes_[leaf].right = partner; }
    if (tmp >= 0) {
        const int pp = nodes_[partner].parent;
        nodes_[tmp].parent = pp;
        if (nodes_[pp].left == partner) nodes_[pp].left = tmp;
        else if (nodes_[pp].right == partner) nodes_[pp].right = tmp;
        nodes_[partner].parent = leaf;
    }
}

void tree::nni(unsigned int branch_no, tree& L, tree& R) const {
    L.nodes_ = nodes_;
    L.root_ = root_;
    L.clades.clear();
    R.nodes_ = nodes_;
    R.root_ = root_;
    R.clades.clear();
    const int v = internal_branch_by_number(branch_no);
    if (branch_no == 1 && (n_sub_tips(v) == 1 || n_sub_tips(root_) - n_sub_tips(v) == 1)) {
        L.clear();
        R.clear();
        return;
    }
    L.interchange(v, true);
    R.interchange(v, false);
}

}  // namespace phylo_amd
