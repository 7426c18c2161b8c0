// ref_harness.cpp -- TEST INFRASTRUCTURE ONLY.
//
// Drives the reference's own likelihood objects (simpleML, sub_model,
// rate_model, tree, matrix_parser; compiled from the reference sources by
// oracle/Makefile into oracle/_ref/) through their public interfaces, to write
// golden vectors and to time the reference's CPU path.  Nothing here is part
// of the product; it is never shipped or measured as the GPU path.
//
// Modes (all numbers printed with %.17g):
//   persite  TREE FASTA ALPHA TR SPEC PARAMS FREQS  per-site lnL (one tree per site,
//            set_char(0, ...) as treeator.cpp:723-724 does), then the total of one
//            multi-site evaluation (set_char(s, ...) for every site)
//   ratelik  TREE FASTA ALPHA TR SPEC PARAMS FREQS TIMEARG CLADEARG
//            multi-site total with a rate_model (treeator.cpp:919, :923)
//   pmatrix  S TR SPEC PARAMS FREQS T1 [T2 ...]     sub_model::set_P_matrix dumps
//   tmbl     TIMEARG B1 D1 [B2 D2 ...]              get_time_mod_branch_length dumps
//   ratearg  clade|time ARG                        the rate_model pars_clade_rates_arg /
//            pars_rates_and_times_arg build, through its getters
//   matrix   FORMAT FILE                            rows the reference's file_parser + matrix_parser
//            read (fasta | phylip | nexus, DNA alphabet), state sets as integers
//   ntrees   FILE                                   Translate table and trees of a NEXUS TREES block
//   nexus    TREE                                   print_nexus_tree_intro / print_tree_to_nexus /
//            print_nexus_tree_end, empty then filled Translate table
//   parsex   TREE                                   print_newick of trees read with "[&...]" kept
//   nodes    TREE FASTA ALPHA TR SPEC PARAMS FREQS NSITES
//            draw_normalized_likelihood_on_nodes + print_newick
//   parse    TREE                                   print_newick of each parsed tree
//   nni      TREE                                   treebender --nni all tree list
//   time     TREE FASTA ALPHA TR SPEC PARAMS FREQS REPS
//            seconds per multi-site calculate_log_likelihood call
// ALPHA is dna | protein | binary | a whitespace alphabet file; SPEC/PARAMS/FREQS
// are comma lists ("-" for none).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "simpleML.h"
#include "file_parser.h"
#include "matrix_parser.h"
#include "rate_model.h"
#include "sub_model.h"

using namespace std;

static vector<double> dlist(const string& s) {
    vector<double> v;
    if (s == "-") return v;
    stringstream ss(s);
    string tok;
    while (getline(ss, tok, ',')) v.push_back(atof(tok.c_str()));
    return v;
}

static string slurp(const char* path) {
    ifstream f(path);
    stringstream ss;
    ss << f.rdbuf();
    return ss.str();
}

static void load_alphabet(const string& name, map<char, bitset<SIZE> >& alphabet) {
    alphabet.clear();
    if (name == "dna") alphabet::set_alphabet_dna(alphabet);
    else if (name == "protein") alphabet::set_alphabet_amino_acid(alphabet);
    else if (name == "binary") alphabet::set_alphabet_binary(alphabet);
    else {
        ifstream f(name.c_str());
        alphabet_parser p(f, alphabet);
        p.pars();
    }
}

// Data preparation as treeator.cpp:395-466 and :549-636 do it.
struct Prepared {
    vector<character_vector> chars;
    map<char, bitset<SIZE> > alphabet;
    unsigned int n_states;
};

static void prepare(const char* fasta, const string& alpha, Prepared& P) {
    load_alphabet(alpha, P.alphabet);
    bitset<SIZE> all;
    for (map<char, bitset<SIZE> >::const_iterator i = P.alphabet.begin(); i != P.alphabet.end(); ++i) all |= i->second;
    partitions regions;
    regions.add_alphabet("first", P.alphabet);
    regions.add_partition(0, 0, "default", "first");
    ifstream f(fasta);
    matrix_parser parser(f, P.chars, regions, "fasta");
    parser.pars();
    for (vector<character_vector>::iterator i = P.chars.begin(); i != P.chars.end(); ++i) i->set_empty_to(all);
    P.n_states = 0;
    for (vector<character_vector>::iterator i = P.chars.begin(); i != P.chars.end(); ++i)
        if (i->highest_char_state() + 1 > P.n_states) P.n_states = i->highest_char_state() + 1;
}

static bool setup_model(sub_model& model, const string& spec, const string& params, const string& freqs) {
    vector<double> sp = dlist(spec), pa = dlist(params), fr = dlist(freqs);
    for (unsigned int i = 0; i < sp.size(); ++i)
        if (!model.set_rate_spec(i, (unsigned int)sp[i])) return false;
    if (pa.size() != model.get_n_rates()) return false;
    for (unsigned int i = 0; i < pa.size(); ++i)
        if (!model.set_rate(i, pa[i])) return false;
    if (model.is_tr()) {
        if (fr.empty()) model.set_freq_same();
        else
            for (unsigned int i = 0; i < fr.size(); ++i) model.set_freq(i, fr[i]);
    }
    return true;
}

int main(int argc, char** argv) {
    if (argc < 3) { cerr << "usage: see header" << endl; return 2; }
    string mode = argv[1];
    if (mode == "persite" || mode == "ratelik" || mode == "nodes" || mode == "time") {
        if (argc < 9) return 2;
        Prepared D;
        prepare(argv[3], argv[4], D);
        bool tr = atoi(argv[5]) != 0;
        sub_model model(D.n_states, tr);
        if (!setup_model(model, argv[6], argv[7], argv[8])) { cerr << "model setup failed" << endl; return 3; }
        string tree_text = slurp(argv[2]);
        unsigned int n_char = D.chars.empty() ? 0 : D.chars.begin()->n_char();
        if (mode == "persite") {
            for (unsigned int s = 0; s < n_char; ++s) {
                stringstream ts(tree_text);
                simpleML t;
                t.tree_file_parser(ts);
                t.init();
                for (vector<character_vector>::iterator i = D.chars.begin(); i != D.chars.end(); ++i)
                    t.set_char(0, i->get_taxon(), i->get_character(s), model);
                printf("%.17g\n", t.calculate_log_likelihood(model));
            }
            stringstream ts(tree_text);
            simpleML t;
            t.tree_file_parser(ts);
            t.init();
            for (unsigned int s = 0; s < n_char; ++s)
                for (vector<character_vector>::iterator i = D.chars.begin(); i != D.chars.end(); ++i)
                    t.set_char(s, i->get_taxon(), i->get_character(s), model);
            printf("total %.17g\n", t.calculate_log_likelihood(model));
        } else if (mode == "ratelik") {
            if (argc < 11) return 2;
            rate_model rate_mod;
            vector<string> taxon_sets;
            string targ = argv[9], carg = argv[10];
            if (targ != "-" && !rate_mod.pars_rates_and_times_arg(targ.c_str())) return 4;
            if (carg != "-" && !rate_mod.pars_clade_rates_arg(carg.c_str(), taxon_sets)) return 4;
            stringstream ts(tree_text);
            simpleML t;
            t.tree_file_parser(ts);
            if (!taxon_sets.empty()) t.add_taxon_sets(taxon_sets);
            t.init();
            for (unsigned int s = 0; s < n_char; ++s)
                for (vector<character_vector>::iterator i = D.chars.begin(); i != D.chars.end(); ++i)
                    t.set_char(s, i->get_taxon(), i->get_character(s), model);
            double l = carg != "-" ? t.calculate_likelihood_rate_change_at_nodes(rate_mod, model)
                                   : t.calculate_likelihood_rate_change_in_time(rate_mod, model);
            printf("total %.17g\n", l);
        } else if (mode == "nodes") {
            if (argc < 10) return 2;
            unsigned int ns = atoi(argv[9]);
            if (ns > n_char) ns = n_char;
            stringstream ts(tree_text);
            simpleML t;
            t.tree_file_parser(ts);
            t.init();
            for (unsigned int s = 0; s < ns; ++s)
                for (vector<character_vector>::iterator i = D.chars.begin(); i != D.chars.end(); ++i)
                    t.set_char(s, i->get_taxon(), i->get_character(s), model);
            printf("total %.17g\n", t.calculate_log_likelihood(model));
            fflush(stdout);
            t.draw_normalized_likelihood_on_nodes();
            t.print_newick();
        } else {
            if (argc < 10) return 2;
            int reps = atoi(argv[9]);
            stringstream ts(tree_text);
            simpleML t;
            t.tree_file_parser(ts);
            t.init();
            for (unsigned int s = 0; s < n_char; ++s)
                for (vector<character_vector>::iterator i = D.chars.begin(); i != D.chars.end(); ++i)
                    t.set_char(s, i->get_taxon(), i->get_character(s), model);
            double best = 1e300, l = 0;
            for (int r = 0; r < reps; ++r) {
                auto a = chrono::steady_clock::now();
                l = t.calculate_log_likelihood(model);
                auto b = chrono::steady_clock::now();
                double sec = chrono::duration<double>(b - a).count();
                if (sec < best) best = sec;
            }
            printf("total %.17g\nseconds %.9g\nsites %u\nstates %u\n", l, best, n_char, D.n_states);
        }
        return 0;
    }
    if (mode == "pmatrix") {
        if (argc < 8) return 2;
        unsigned int S = atoi(argv[2]);
        sub_model model(S, atoi(argv[3]) != 0);
        if (!setup_model(model, argv[4], argv[5], argv[6])) return 3;
        for (int a = 7; a < argc; ++a) {
            model.set_P_matrix(atof(argv[a]));
            for (unsigned int i = 0; i < S; ++i)
                for (unsigned int j = 0; j < S; ++j) printf("%.17g%c", model.get_P_value(i, j), j + 1 < S ? ' ' : '\n');
        }
        return 0;
    }
    if (mode == "matrix") {
        // the reference's own readers as treeator drives them (treeator.cpp:438-455)
        std::ifstream f(argv[3]);
        file_parser input(&f);
        input.set_file_type(argv[2]);
        if (input.test_file_type("nexus") && !input.move_to_next_X_block(nexus_block::DATA)) return 5;
        map<char, bitset<SIZE> > alpha;
        alphabet::set_alphabet_dna(alpha);
        partitions regions;
        regions.add_alphabet("first", alpha);
        regions.add_partition(0, 0, "default", "first");
        vector<character_vector> rows;
        matrix_parser parser(*input.file_stream, rows, regions, input.get_file_type());
        parser.pars();
        for (auto& r : rows) {
            printf(">%s\n", r.get_taxon().c_str());
            for (unsigned int k = 0; k < r.n_char(); ++k) printf("%s%lu", k ? "," : "", r.get_character(k).to_ulong());
            printf("\n");
        }
        return 0;
    }
    if (mode == "ntrees") {
        // NEXUS TREES block: Translate table, then every tree (treeator.cpp:490-530)
        std::ifstream f(argv[2]);
        file_parser input(&f);
        input.set_file_type("nexus");
        if (!input.move_to_next_X_block(nexus_block::TREES)) return 5;
        map<string, string> taxa_trans;
        char cmd = input.read_next_nexus_command();
        if (cmd == nexus_command::TRANSLATE) {
            input.read_translate_parameters(taxa_trans);
            cmd = input.read_next_nexus_command();
        }
        for (auto& kv : taxa_trans) printf("translate [%s] [%s]\n", kv.first.c_str(), kv.second.c_str());
        unsigned int n = 0;
        while (true) {
            if (n != 0) cmd = input.read_next_nexus_command();
            if (!(cmd == nexus_command::TREE && input.move_to_start_of_tree())) break;
            tree t;
            t.tree_file_parser(*input.file_stream, taxa_trans, false);
            if (t.empty()) break;
            ++n;
            t.print_newick();
        }
        return 0;
    }
    if (mode == "nexus" || mode == "parsex") {
        // NEXUS writer (an empty Translate table, filled by the intro, then
        // the filled one), or newick of trees read with "[&...]" annotations
        std::stringstream ts(slurp(argv[2]));
        std::map<std::string, std::string> no_translation;
        while (true) {
            tree t;
            t.tree_file_parser(ts, no_translation, mode == "parsex");
            if (t.empty()) break;
            if (mode == "parsex") { t.print_newick(); continue; }
            std::map<std::string, std::string> tt;
            std::string name = "tree1";
            t.print_nexus_tree_intro(tt);
            t.print_tree_to_nexus(name, true, true, tt);
            t.print_nexus_tree_end();
            t.print_nexus_tree_intro(tt);
            t.print_tree_to_nexus(name, false, true, tt);
            t.print_nexus_tree_end();
        }
        return 0;
    }
    if (mode == "ratearg") {
        // the model a -V (clade) or time-rate argument builds, through the getters
        rate_model rate_mod;
        const std::string kind = argv[2];
        std::vector<std::string> sets;
        const bool ok = kind == "clade" ? rate_mod.pars_clade_rates_arg(argv[3], sets)
                                        : rate_mod.pars_rates_and_times_arg(argv[3]);
        printf("ok %d\n", ok ? 1 : 0);
        printf("time_rates %u cut_offs %u\n", rate_mod.get_n_time_rates(), rate_mod.get_n_cut_offs());
        for (unsigned int i = 0; i < rate_mod.get_n_cut_offs(); ++i)
            printf("cut %.17g rate %.17g\n", (double)rate_mod.get_cut_off(i), (double)rate_mod.get_rate_in_time(i));
        printf("clade_rates %u clades %u\n", rate_mod.get_n_clade_rates(), rate_mod.get_n_clades());
        for (unsigned int i = 0; i < rate_mod.get_n_clades(); ++i)
            printf("clade %u rate %.17g\n", i, (double)rate_mod.get_rate_clade(i));
        for (size_t i = 0; i < sets.size(); ++i) printf("set %zu [%s]\n", sets[i].size(), sets[i].c_str());
        return 0;
    }
    if (mode == "tmbl") {
        rate_model rate_mod;
        string targ = argv[2];
        if (targ != "-" && !rate_mod.pars_rates_and_times_arg(targ.c_str())) return 4;
        for (int a = 3; a + 1 < argc; a += 2)
            printf("%.17g\n", rate_mod.get_time_mod_branch_length(atof(argv[a]), (float)atof(argv[a + 1])));
        return 0;
    }
    if (mode == "parse" || mode == "nni") {
        string text = slurp(argv[2]);
        stringstream ts(text);
        while (true) {
            tree t;
            t.tree_file_parser(ts);
            if (t.empty()) break;
            if (mode == "parse") { t.print_newick(); continue; }
            vector<tree> out;
            out.push_back(t);
            unsigned int n_tips = t.n_tips();
            if (n_tips >= 4)
                for (unsigned int i = 2; i <= n_tips - 3; ++i) {
                    tree one, two;
                    out[0].nni(i, one, two);
                    out.push_back(one);
                    out.push_back(two);
                }
            for (size_t k = 0; k < out.size(); ++k) out[k].print_newick();
            break;
        }
        return 0;
    }
    cerr << "unknown mode" << endl;
    return 2;
}
