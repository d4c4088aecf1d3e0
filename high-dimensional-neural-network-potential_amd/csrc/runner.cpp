// RuNNer parameter parser: the semantics of parseRuNNerInput
// (src/potentials/hdnnps.f90:34-233) and its helpers, with the reference's
// print+stop failures turned into hdnnp::Error.
#include "runner.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <numeric>
#include <sstream>

namespace hdnnp {

static const char *kSymbols[118] = {
    "H",  "He", "Li", "Be", "B",  "C",  "N",  "O",  "F",  "Ne", "Na", "Mg", "Al", "Si", "P",
    "S",  "Cl", "Ar", "K",  "Ca", "Sc", "Ti", "V",  "Cr", "Mn", "Fe", "Co", "Ni", "Cu", "Zn",
    "Ga", "Ge", "As", "Se", "Br", "Kr", "Rb", "Sr", "Y",  "Zr", "Nb", "Mo", "Tc", "Ru", "Rh",
    "Pd", "Ag", "Cd", "In", "Sn", "Sb", "Te", "I",  "Xe", "Cs", "Ba", "La", "Ce", "Pr", "Nd",
    "Pm", "Sm", "Eu", "Gd", "Tb", "Dy", "Ho", "Er", "Tm", "Yb", "Lu", "Hf", "Ta", "W",  "Re",
    "Os", "Ir", "Pt", "Au", "Hg", "Tl", "Pb", "Bi", "Po", "At", "Rn", "Fr", "Ra", "Ac", "Th",
    "Pa", "U",  "Np", "Pu", "Am", "Cm", "Bk", "Cf", "Es", "Fm", "Md", "No", "Lr", "Rf", "Db",
    "Sg", "Bh", "Hs", "Mt", "Ds", "Rg", "Cn", "Nh", "Fl", "Mc", "Lv", "Ts", "Og"};

// elemSymToNum (constants.f90:95-111) applied to a character(len=2) value.
int elem_sym_to_num(const std::string &sym)
{
    std::string s = sym.substr(0, 2);
    while (!s.empty() && s.back() == ' ') s.pop_back();
    while (!s.empty() && s.front() == ' ') s.erase(s.begin());
    for (int i = 0; i < 118; ++i)
        if (s == kSymbols[i]) return i + 1;
    return -1;
}

// splitString (util.f90:128-161): words are separated by blanks only.
static std::vector<std::string> split_blanks(const std::string &line)
{
    std::vector<std::string> w;
    size_t i = 0;
    while (i < line.size()) {
        while (i < line.size() && line[i] == ' ') ++i;
        size_t j = i;
        while (j < line.size() && line[j] != ' ') ++j;
        if (j > i) w.push_back(line.substr(i, j - i));
        i = j;
    }
    return w;
}

static double read_real(const std::string &tok, const std::string &what)
{
    std::string t = tok;
    for (char &c : t)
        if (c == 'd' || c == 'D') c = 'e';
    char *end = nullptr;
    const double v = std::strtod(t.c_str(), &end);
    if (end == t.c_str() || (*end != '\0' && *end != ',')) throw Error(2, "bad real '" + tok + "' in " + what);
    return v;
}

static int read_int(const std::string &tok, const std::string &what)
{
    char *end = nullptr;
    const long v = std::strtol(tok.c_str(), &end, 10);
    if (end == tok.c_str() || (*end != '\0' && *end != ',')) throw Error(2, "bad integer '" + tok + "' in " + what);
    return (int)v;
}

static std::string fmt_f(double x, int w, int d)
{
    char buf[64];
    std::snprintf(buf, sizeof buf, "%*.*f", w, d, x);
    std::string s(buf);
    return (int)s.size() > w ? std::string(w, '*') : s;
}

static std::string fmt_i(int n, int w)
{
    char buf[32];
    std::snprintf(buf, sizeof buf, "%*d", w, n);
    std::string s(buf);
    return (int)s.size() > w ? std::string(w, '*') : s;
}

// symfunctionstring (hdnnps.f90:254-267), blank-padded to character(len=256).
static std::string sf_key(const SymFn &s)
{
    std::string k;
    if (s.type == 2)
        k = fmt_i(2, 2) + fmt_f(s.rcut, 9, 5) + fmt_f(s.eta, 9, 5) + fmt_f(s.rs, 9, 5) + fmt_i(s.e1, 3);
    else   // types 3 and 9 (RuNNer's wide angular function, in type 3's form)
        k = fmt_i(s.type, 2) + fmt_f(s.rcut, 9, 5) + fmt_f(s.eta, 9, 5) + fmt_f(s.zeta, 9, 5) +
            fmt_f(s.lambda + 10, 9, 5) + fmt_i(s.e1, 3) + fmt_i(s.e2, 3);
    k.resize(256, ' ');
    return k;
}

// Reader for list-directed records: a READ consumes values from as many lines
// as it needs, and the next READ starts on the following line.
class RecordReader {
  public:
    explicit RecordReader(const std::string &path) : path_(path)
    {
        std::ifstream f(path);
        if (!f) throw Error(2, path + " not found");
        std::string ln;
        while (std::getline(f, ln)) lines_.push_back(ln);
    }
    std::vector<std::string> record(size_t n)
    {
        std::vector<std::string> out;
        while (out.size() < n) {
            if (cur_ >= lines_.size()) throw Error(2, "unexpected end of " + path_);
            std::string ln = lines_[cur_++];
            for (char &c : ln)
                if (c == '\t' || c == ',' || c == '\r') c = ' ';
            for (auto &w : split_blanks(ln)) {
                if (out.size() < n) out.push_back(w);
            }
        }
        return out;
    }

  private:
    std::string path_;
    std::vector<std::string> lines_;
    size_t cur_ = 0;
};

HostModel read_runner(const std::string &dir_in)
{
    std::string dir = dir_in;
    if (!dir.empty() && dir.back() != '/') dir += '/';
    std::ifstream in(dir + "input.nn");
    if (!in) throw Error(2, dir + "input.nn not found");

    HostModel m;
    int nels = -1, nlayers = -1;
    std::vector<std::string> elements;
    std::vector<int> nnodes;
    std::vector<char> acts;
    std::vector<std::pair<int, double>> atom_energy;
    std::vector<std::vector<SymFn>> lists(119);
    std::vector<bool> allocated(119, false);
    double maxrc = 0.0;

    std::string raw;
    while (std::getline(in, raw)) {                  // :66-198
        std::string line = raw.substr(0, 256);
        if (!line.empty() && line.back() == '\r') line.pop_back();
        auto w = split_blanks(line);
        if (w.empty()) continue;
        const std::string &key = w[0];
        auto need = [&](size_t n) {
            if (w.size() < n) throw Error(2, "too few values on input.nn line: " + line);
        };
        if (key == "number_of_elements") {
            need(2);
            nels = read_int(w[1], "number_of_elements");
            if (nels < 1 || nels > kMaxElements)
                throw Error(6, "number_of_elements outside 1.." + std::to_string(kMaxElements));
        } else if (key == "elements") {
            if (nels < 0) throw Error(2, "'elements' before 'number_of_elements'");
            need(1 + (size_t)nels);
            elements.clear();
            for (int i = 0; i < nels; ++i) {
                const std::string e = w[1 + i].substr(0, 2);
                const int z = elem_sym_to_num(e);
                if (z < 0) throw Error(2, "unknown element " + e);
                elements.push_back(e);
                allocated[z] = true;
                lists[z].clear();
            }
            std::vector<int> idx(nels);
            std::iota(idx.begin(), idx.end(), 0);
            std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) {
                return elem_sym_to_num(elements[a]) < elem_sym_to_num(elements[b]);
            });
            std::vector<std::string> sorted;
            for (int i : idx) sorted.push_back(elements[i]);
            elements = sorted;
        } else if (key == "atom_energy") {
            need(3);
            const int z = elem_sym_to_num(w[1].substr(0, 2));
            if (z < 0) throw Error(2, "unknown element in atom_energy: " + w[1]);
            atom_energy.emplace_back(z, read_real(w[2], "atom_energy"));
        } else if (key == "scale_symmetry_functions") {
            m.scalesf = true;
        } else if (key == "center_symmetry_functions") {
            m.centersf = true;
        } else if (key == "cutoff_type") {                // not read by the reference
            need(2);
            m.cutoff_type = read_int(w[1], "cutoff_type");
            if (m.cutoff_type < 0 || m.cutoff_type > 8)
                throw Error(3, "cutoff_type " + w[1] + " is not one of RuNNer's cutoff functions 0..8");
            if (w.size() > 2 && read_real(w[2], "cutoff_type") != 0.0)
                throw Error(3, "cutoff_type with an inner cutoff (alpha != 0) is not supported");
        } else if (key == "scale_min_short") {            // not read by the reference
            need(2);
            m.scale_min = read_real(w[1], "scale_min_short");
            m.scale_range = true;
        } else if (key == "scale_max_short") {            // not read by the reference
            need(2);
            m.scale_max = read_real(w[1], "scale_max_short");
            m.scale_range = true;
        } else if (key == "global_hidden_layers_short") {
            need(2);
            nlayers = read_int(w[1], "global_hidden_layers_short");
            if (nlayers < 0 || nlayers + 1 > kMaxLayers)
                throw Error(6, "too many hidden layers for this build");
        } else if (key == "global_nodes_short") {
            if (nlayers < 0) throw Error(2, "'global_nodes_short' before 'global_hidden_layers_short'");
            need(1 + (size_t)nlayers);
            nnodes.assign(nlayers + 1, 1);
            for (int i = 0; i < nlayers; ++i) {
                nnodes[i] = read_int(w[1 + i], "global_nodes_short");
                if (nnodes[i] < 1 || nnodes[i] > kMaxNodes)
                    throw Error(6, "hidden layer width outside 1.." + std::to_string(kMaxNodes));
            }
        } else if (key == "global_activation_short") {
            if (nlayers < 0) throw Error(2, "'global_activation_short' before 'global_hidden_layers_short'");
            need(2 + (size_t)nlayers);
            acts.assign(nlayers + 1, 'l');
            for (int i = 0; i <= nlayers; ++i) {
                acts[i] = w[1 + i][0];
                if (acts[i] != 'l' && acts[i] != 't' && acts[i] != 'p' && acts[i] != 'r')
                    throw Error(2, "Invalid activation function!");   // neuralNetworks.f90:52
            }
        } else if (key == "symfunction_short") {
            need(3);
            const int name = read_int(w[2], "symfunction_short");
            SymFn s;
            const int centre = elem_sym_to_num(w[1].substr(0, 2));
            if (name == 2) {
                need(7);
                s.type = 2;
                s.e1 = elem_sym_to_num(w[3].substr(0, 2));
                s.eta = read_real(w[4], "symfunction_short");
                s.rs = read_real(w[5], "symfunction_short");
                s.rcut = read_real(w[6], "symfunction_short");
            } else if (name == 3 || name == 9) {   // 9: RuNNer's wide angular G5 (the reference reads 2 and 3)
                need(9);
                s.type = name;
                s.e1 = elem_sym_to_num(w[3].substr(0, 2));
                s.e2 = elem_sym_to_num(w[4].substr(0, 2));
                s.eta = read_real(w[5], "symfunction_short");
                s.lambda = read_real(w[6], "symfunction_short");
                s.zeta = read_real(w[7], "symfunction_short");
                s.rcut = read_real(w[8], "symfunction_short");
            } else {
                // the reference skips other types silently, and its weights
                // would then be read with the wrong layout: refused here
                throw Error(3, "symmetry-function type " + w[2] + " is not supported (types 2, 3 and 9 are)");
            }
            if (s.rcut > maxrc) maxrc = s.rcut;
            if (centre < 0 || !allocated[centre])
                throw Error(2, "symmetry function for element '" + w[1] + "' that is not in 'elements'");
            if (s.e1 < 0 || (s.type != 2 && s.e2 < 0))
                throw Error(2, "unknown neighbour element in: " + line);
            if ((int)lists[centre].size() + 1 > kMaxSF)
                throw Error(2, "Error: more symmetry functions in input file than nmaxSFs allows");
            lists[centre].push_back(s);
        }
    }
    if (elements.empty()) throw Error(2, "input.nn defines no elements");
    if (nlayers < 0 || nnodes.empty() || acts.empty())
        throw Error(2, "input.nn lacks global_hidden_layers_short/global_nodes_short/global_activation_short");
    m.maxrc = maxrc;
    if (m.scale_range && !(m.scale_max > m.scale_min))
        throw Error(3, "scale_max_short must be greater than scale_min_short");

    for (auto &e : elements) {                        // :204-212, sortSymFunctions :235-252
        ElementModel em;
        em.Z = elem_sym_to_num(e);
        auto &lst = lists[em.Z];
        std::vector<std::string> keys;
        for (auto &s : lst) keys.push_back(sf_key(s));
        std::vector<int> idx(lst.size());
        std::iota(idx.begin(), idx.end(), 0);
        std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return keys[a] < keys[b]; });
        for (int i : idx) em.sfs.push_back(lst[i]);
        for (auto &ae : atom_energy)
            if (ae.first == em.Z) em.atom_energy = ae.second;
        // preconditions of calcAllSymmetryFunctions (symmetryFunctions.f90:183-199)
        if (em.sfs.empty()) {
            em.valid = false;
            em.why = "element " + e + " has no symmetry functions";
        } else {
            for (auto &s : em.sfs) {
                const double dr = s.rcut - em.sfs[0].rcut;
                if (dr * dr > 1.e-10) {
                    em.valid = false;
                    em.why = "Error: expecting same rc for all symfunctions in calcAllSymmetryFunctions";
                }
            }
        }
        m.el.push_back(std::move(em));
    }

    {   // readScaling, hdnnps.f90:269-300
        RecordReader rr(dir + "scaling.data");
        for (size_t i = 0; i < m.el.size(); ++i) {
            for (size_t s = 0; s < m.el[i].sfs.size(); ++s) {
                auto r = rr.record(5);
                const int riel = read_int(r[0], "scaling.data"), risf = read_int(r[1], "scaling.data");
                auto &sf = m.el[i].sfs[s];
                sf.sfmin = read_real(r[2], "scaling.data");
                sf.sfmax = read_real(r[3], "scaling.data");
                sf.sfmean = read_real(r[4], "scaling.data");
                if (riel != (int)i + 1 || risf != (int)s + 1)
                    throw Error(2, "Error: wrong index reading scaling.data");
            }
        }
        auto r = rr.record(2);
        m.emin = read_real(r[0], "scaling.data");
        m.emax = read_real(r[1], "scaling.data");
    }

    for (auto &em : m.el) {                           // nn_new + readWeights :216-229, 302-330
        char fn[32];
        std::snprintf(fn, sizeof fn, "weights.%03d.data", em.Z);
        RecordReader rr(dir + fn);
        int nin = (int)em.sfs.size();
        for (size_t l = 0; l < nnodes.size(); ++l) {
            Layer L;
            L.nin = nin;
            L.nout = nnodes[l];
            L.act = acts[l];
            L.W.resize((size_t)L.nin * L.nout);
            L.b.resize(L.nout);
            for (int j = 0; j < L.nin; ++j)
                for (int i = 0; i < L.nout; ++i) L.W[i + (size_t)L.nout * j] = read_real(rr.record(1)[0], fn);
            for (int i = 0; i < L.nout; ++i) L.b[i] = read_real(rr.record(1)[0], fn);
            em.layers.push_back(std::move(L));
            nin = nnodes[l];
        }
    }
    return m;
}

}  // namespace hdnnp
