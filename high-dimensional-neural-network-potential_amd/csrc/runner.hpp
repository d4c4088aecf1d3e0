// RuNNer parameter files -> host model (parseRuNNerInput semantics).
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

namespace hdnnp {

constexpr int kMaxElements = 16;   // elements per model in this build
constexpr int kMaxSF = 200;        // nmaxSFs, hdnnps.f90:48
constexpr int kMaxLayers = 8;
constexpr int kMaxNodes = 256;     // widest MLP layer in this build

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

struct SymFn {
    int type = 0;          // 2 (G2, radial), 3 (G4, angular) or 9 (G5, wide angular: no fc(r_jk))
    int e1 = 0, e2 = 0;    // neighbour element(s) as Z, as written in input.nn
    double rcut = 0, eta = 0, rs = 0, lambda = 0, zeta = 0;
    double sfmin = 0, sfmax = 0, sfmean = 0;
};

struct Layer {
    int nin = 0, nout = 0;
    char act = 'l';
    std::vector<double> W;   // W(i,j) at [i + nout*j] (reference column-major)
    std::vector<double> b;
};

struct ElementModel {
    int Z = 0;
    double atom_energy = 0;
    std::vector<SymFn> sfs;      // sorted as sortSymFunctions leaves them
    std::vector<Layer> layers;
    bool valid = true;           // calcAllSymmetryFunctions preconditions hold
    std::string why;             // reason when !valid
};

struct HostModel {
    std::vector<ElementModel> el;   // ascending Z
    bool scalesf = false, centersf = false;
    double maxrc = 0, emin = 0, emax = 0;
    // RuNNer keywords beyond the reference's: cutoff_type (0..8; 2, the
    // reference's tanh^3, when absent) and scale_min_short / scale_max_short
    // (RuNNer's min-max scaling of the SFs to [scale_min, scale_max])
    int cutoff_type = 2;
    bool scale_range = false;
    double scale_min = 0, scale_max = 1;
    int rank_of_Z(int z) const {
        for (size_t e = 0; e < el.size(); ++e)
            if (el[e].Z == z) return (int)e;
        return -1;
    }
};

int elem_sym_to_num(const std::string &sym);
HostModel read_runner(const std::string &dir);

}  // namespace hdnnp
