// Device-side data structures and kernel launchers of the HDNNP engine.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hdnnp {

constexpr int kMaxEl = 16;       // elements per model (= kMaxElements)
constexpr int kTripBlk = 64;   // ints per atom of Work::tblk
constexpr int kMaxLay = 8;       // weight layers
constexpr int kNonIntegralZeta = -(1 << 30);   // g4_ipow marker

// Per-element scalars, kept in device memory: kernels index them with the
// (wave-uniform) element rank, which compiles to scalar loads. (Indexing an
// array member of the by-value DevModel argument with a runtime index would
// make the compiler copy the argument into private scratch memory.)
struct ElemInfo {
    int nsf, sf_off;
    int g2_start, g2_count, g4_start, g4_count;
    double rc, atom_energy;
};

// One symmetry function of the per-function path and of the G5 kernels:
// type 2 (G2, neighbour element rank f), 3 (G4) or 9 (G5, element pair
// {f, g}); s its index in its element's sorted list; pre = 2^(1 - zeta).
struct SfParam {
    int type, f, g, s;
    double eta, rs, lambda, zeta, pre;
};

// Model tables in device memory; the struct itself is a kernel argument.
// G2 entries of centre element e with neighbour element f live at
// g2_tab[e*kMaxEl+f] = (start, count); G4 entries of centre e with the
// unordered neighbour pair (f <= g) at g4_tab[(e*kMaxEl+f)*kMaxEl+g].
// Within a G4 block the entries keep the reference's sorted order, so equal
// eta are adjacent; g4_neweta marks where exp(-eta*sum r^2) must be redone.
struct DevModel {
    int nel;
    int nsfmax;
    int nsf[kMaxEl];
    int sf_off[kMaxEl];
    double rc[kMaxEl];
    const int2 *g2_tab;
    const int2 *g4_tab;
    const int *g2_s;
    const double *g2_eta, *g2_rs;
    const int *g4_s;
    const double *g4_eta, *g4_lam, *g4_zeta, *g4_pre;
    const int *g4_ipow;      // zeta-1 when zeta is integral, else kNonIntegralZeta
    const int *g4_sq;        // k when zeta = 2^k (k >= 0), else -1
    const int *g4_neweta;
    const int *g2_f;         // neighbour element rank of each G2 entry
    // FP32 copies of the tables (the mixed path reads these, no conversions)
    const float *g2f_eta, *g2f_rs, *g4f_eta, *g4f_lam, *g4f_zeta, *g4f_pre;
    const float *g2f_ea, *g4f_ea;   // -eta * log2(e): exp(-eta x) = exp2(ea x)
    // the entries of centre element e are contiguous: [start, start+count)
    int g2_start[kMaxEl], g2_count[kMaxEl];
    int g4_start[kMaxEl], g4_count[kMaxEl];
    int g4_max;              // largest g4_count
    int g4_lampm;            // 1 when every G4 lambda is +1 or -1 and every zeta >= 1
    int g4_nblk;             // most element-pair blocks with G4 functions of one centre element
    int powbits;             // bits of the largest |g4_ipow| of integral zeta
    const ElemInfo *einfo;   // [nel]
    const double *sf_mean;   // [sf_off[e] + s]: scaled G = (G - sf_mean) / sf_den
    const double *sf_den;
    int nlayers;
    int nodes[kMaxLay];
    char act[kMaxLay];
    const double *W64[kMaxEl];   // per layer: W(o,j) at [j*nout + o], then bias[nout]
    const float *W32[kMaxEl];
    double atom_energy[kMaxEl];
    const int *zlut;             // Z (0..127) -> element rank; -1 absent, -2-rank unusable
    int fallback_el;             // usable element given to flagged atoms (results discarded)
    int cutoff_type;             // RuNNer cutoff_type 0..8 (2: tanh^3, the reference's)
    // G5 (RuNNer type 9) functions of centre element e: g5[g5_start[e] ..
    // + g5_count[e]); g5_total = 0 when the model has none
    const SfParam *g5;
    int g5_start[kMaxEl], g5_count[kMaxEl];
    int g5_total;
};

// Cell grid of one structure of an evaluation (host-computed, reference
// formulas). The structures of a batch are concatenated: atoms
// [atom_off, atom_off + nat), cells [cell_off, cell_off + ng0*ng1*ng2).
struct StructGeo {
    double lat[9];      // list-frame lattice, lat[3k+a] = component a of vector k
    double inv[9];      // its inverse (inv3DM form)
    double origin[3];   // artificial-box origin (nonperiodic), else 0
    int periodic;
    int ng[3];          // grid cells per lattice vector
    int nc[3];          // stencil half-widths (getNcells of the grid)
    int S;              // stencil cells, (2nc0+1)(2nc1+1)(2nc2+1)
    int atom_off, nat;
    int cell_off, ncell;
    long long ps_off;   // first pos_start row of its atoms
};

// Kernel argument: the structures of one evaluation.
struct Geo {
    const StructGeo *sg;    // [nstruct], device memory
    double rc2;             // (largest cutoff)^2
    int nstruct;
    int ncell;              // cells of all structures
    int nat;                // atoms of all structures
};

// Device work arrays of one evaluation.
struct Work {
    int nat;             // atoms of all structures
    const double *pos;   // [nat][3] caller positions
    int *sid;            // [nat] structure of each atom
    int *csid;           // [ncell] structure of each cell
    const int *Z;        // [nat]
    int *eidx;           // element rank per atom
    double4 *wpos;       // wrapped positions in the list frame
    int *cellkey;        // cell id per atom
    int *cell_atoms;     // atoms sorted by cell
    int *cellkey_sorted;
    int *cell_start, *cell_end;
    // Neighbour entries of atom i occupy slots [i*cap, i*cap + count[i]) of
    // the entry arrays (cap a power of two, fixed for one evaluation), in the
    // reference's list order: stencil cell, then ascending atom index. An
    // atom holds at most lim <= cap entries: lim is the staging size of the
    // symmetry-function kernels (64, 96, 128, 256, 512 or 1024).
    int cap, lim;
    int *count;          // [nat] neighbours per atom (at most lim)
    // G4 triplets of each atom as the value pass compacts them (one atom per
    // wavefront only): packed j | k << 10 | block << 20 over staged
    // neighbours, at most tcap per atom; tcount[i] may exceed tcap, and the
    // gradient pass then compacts that atom's triplets itself. trip = null:
    // not written.
    int *trip, *tcount;
    int tcap;
    // Per atom kTripBlk ints: the element-pair blocks of its triplet list
    // (which the FP32 values pass writes in block order): the count (| 1 <<
    // 16 when the list holds 16-bit entries j | k << 8, one region per
    // block), then start | count << 11 | block << 23 each. null: not written.
    int *tblk;
    int *ent_j;          // [nat*cap] neighbour atom
    int *ent_shift;      // [nat*cap] packed image shift (3 x int8)
    int *ent_s;          // [nat*cap] stencil-cell index of the entry
    int *rev;            // [nat*cap] partner entry in the neighbour's list, -1 if none
    unsigned short *pos_start;   // per atom, S rows (StructGeo::ps_off): first list slot of
                                 // each stencil cell, or null
    int *perm;           // atoms sorted by element
    int *elem_off;       // [kMaxEl+1] start of each element in perm
    int *elem_count;     // [kMaxEl]
    int *flags;          // [8]: 0 unknown Z, 1 neighbours needed when above cap,
                         //      2 image-shift overflow, 3 Z of an unusable element
    double *Ei;          // [nat]
    double *vir;         // [nat][9]
    double *force;       // [nat][3]
    double *result;      // [nstruct][10]: energy, W[9] of each structure
};

// Workgroups are dispatched round-robin over the 8 XCDs (workgroup b runs on
// XCD b % 8), each with its own L2. xcd_block maps b to a logical block so
// that every XCD works through one contiguous range of logical blocks, and
// blocks that share data (neighbouring atoms) share an L2. A bijection on
// [0, nb) for any nb.
__device__ inline int xcd_block(int b, int nb)
{
    const int q = nb >> 3, r = nb & 7;
    const int x = b & 7, k = b >> 3;
    return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + k;
}

__host__ __device__ inline int pack_shift(int x, int y, int z)
{
    return (x & 0xFF) | ((y & 0xFF) << 8) | ((z & 0xFF) << 16);
}
__host__ __device__ inline void unpack_shift(int p, int &x, int &y, int &z)
{
    x = (int)(signed char)(p & 0xFF);
    y = (int)(signed char)((p >> 8) & 0xFF);
    z = (int)(signed char)((p >> 16) & 0xFF);
}

// ---- launchers (kernels_nl.hip) ----
void launch_bbox(const double *pos, const int *off, const int *which, int nwhich, double *out, hipStream_t s);
void launch_nl_build(const Geo &g, Work &w, int ncell, hipStream_t s);
void launch_nl_rev(const Geo &g, Work &w, hipStream_t s);
void launch_fill_int(int *p, int v, long n, hipStream_t s);
void launch_zero3(int *a, long na, int *b, long nb, int *c, long nc, hipStream_t s);
// per-atom set-up of a build: element ids and flags, structure ids of atoms
// and cells, wrapped positions and cell keys, the identity order (iota, or
// null); slot (or null): each atom's arrival slot in its cell, counted in
// cell_end; chunk_cnt (or null): the element sort's per-chunk counts (both
// must be zero)
void launch_prep(const DevModel &m, const Geo &g, Work &w, int *iota, int *slot, int *chunk_cnt, hipStream_t s);
// cell_start / cell_end / cell_atoms from the counts and slots of launch_prep
// (grids of at most kCountSortCells cells, one round of its one-workgroup
// scan; larger grids keep the radix sort, which spreads over the chip;
// tmp: nat ints); its workgroup also turns chunk_cnt into the element sort's
// chunk offsets, elem_off and elem_count
constexpr int kCountSortCells = 24 * 1024;
// (and, in the placement's launch, the element sort's placement by nel elements)
void launch_cell_count_sort(Work &w, int ncell, int nel, const int *slot, int *tmp, int *chunk_cnt, hipStream_t s);
// set-up (as launch_prep, without iota) and cell and element sorts of
// evaluations of at most kSmallSortAtoms atoms in kSmallSortCells cells, in
// one workgroup (cell_atoms, cell_start, cell_end, perm, elem_off,
// elem_count; slot and tmp are nat-int scratch)
constexpr int kSmallSortThreads = 1024, kSmallSortAtoms = 8192, kSmallSortCells = 12288;
// also clears clear[0..nclear) (the evaluation's flags and counters) first
void launch_sort_small(const DevModel &m, const Geo &g, Work &w, int *slot, int *tmp, int *clear, int nclear,
                       hipStream_t s);
// cell_start / cell_end from the key-sorted atoms
void launch_cell_bounds(Work &w, int ncell, hipStream_t s);
// perm (atoms by element, index order within one), elem_off, elem_count;
// chunk_cnt holds kMaxEl ints per elem_sort_chunks(nat) chunk
int elem_sort_chunks(int nat);
void launch_elem_sort(Work &w, int nel, int *chunk_cnt, hipStream_t s);

// ---- launchers (kernels_sf.hip); R = double or float ----
template <typename R>
void launch_sf_values(const DevModel &m, const Geo &g, Work &w, int nmax, int ga, R *G, hipStream_t s);
template <typename R>
void launch_sf_grad(const DevModel &m, const Geo &g, Work &w, int nmax, int ga, const R *dEdG, R *gent,
                    hipStream_t s);
size_t sf_lds_bytes(bool fp64, int nmax, const DevModel &m, bool grad);
// The per-function path (symmetryFunction2 / symmetryFunction3 of the
// reference): for every atom and symmetry function, the centred / scaled
// value G[i * ldG + s] and, when dsf is not null, its gradient rows
// dsf[dsf_off[i] + 3 (s (count_i + 1) + slot)], slot 0 the atom, slot 1 + n
// its list entry n (derivatives with respect to those positions, scaled).
// par[par_off[e] ..] lists element e's functions.
// G5 (RuNNer type 9, no fc(r_jk)): values into G (scaled, row stride
// m.nsfmax), and the gradients of sum_s dE/dG_s G_s added to the per-entry
// rows gent and the atoms' virials (after launch_sf_grad)
template <typename R>
void launch_g5_values(const DevModel &m, const Geo &g, const Work &w, R *G, hipStream_t s);
template <typename R>
void launch_g5_grad(const DevModel &m, const Geo &g, const Work &w, const R *dEdG, R *gent, hipStream_t s);
template <typename R>
void launch_sf_single(const DevModel &m, const Geo &g, const Work &w, const SfParam *par, const int *par_off,
                      const long long *dsf_off, int ldG, double *G, double *dsf, hipStream_t s);
// atoms per wavefront the SF kernels use for a requested tile (1, 2 or 4)
int sf_tile(bool fp64, int nmax, int ga, int nsfmax, int g4max, int nel);
// ---- kernels_mlp.hip: 16 atoms per wavefront, maxw = widest layer/input ----
template <typename R>
void launch_mlp_t(const DevModel &m, Work &w, int maxw, int nblocks, const R *G, R *dEdG, double *Ei, R scale,
                  bool accumulate, hipStream_t s,
                  bool bf16 = false);
size_t mlp_lds_bytes(bool fp64, const DevModel &m, int maxw);
// ---- kernels_misc.hip ----
template <typename R>
void launch_force_gather_t(Work &w, const R *gent, hipStream_t s);
// per-structure sums of Ei and vir: chunks [c.x, c.y) of atoms never cross a
// structure; structure s owns chunks [sc[s], sc[s+1]); partial: 10 per chunk;
// flag_out (or null) receives a copy of w.flags[0..8); done: a counter that
// is zero between launches (the one-launch form for few structures);
// orphans_g (or null): the gradient rows of the force gather, whose
// partnerless entries the first stage's launch adds to w.force
template <typename R>
void launch_reduce(Work &w, const int2 *chunks, int nchunk, const int *sc, int nstruct, double *partial,
                   int *flag_out, unsigned *done, const R *orphans_g, hipStream_t s);

// ---- kernels_md.hip: velocity Verlet and Verlet-skin bookkeeping ----
void launch_md_drift(double *pos, const double *mom, const double *F, double *Fold, const double *mass, double dt,
                     int nat, hipStream_t s);
void launch_md_kick(double *mom, const double *F, const double *Fold, double dt, int nat, hipStream_t s);
void launch_md_disp(const double *pos, const double *posb, int nat, unsigned long long *maxd2, const int *step,
                    hipStream_t s);
void launch_md_refresh(double4 *wpos, const double4 *wposb, const double *pos, const double *posb, int nat,
                       hipStream_t s);
void launch_md_save(double *slot, const double *pos, const double *mom, const double *F, int nat, const int *step,
                    hipStream_t s);
void launch_md_tick(int *step, hipStream_t s);
// a reused-list step's start (drift, displacement into maxd2[*step], wrapped
// positions) and end (kick, state saved after step *step), each one launch
void launch_md_begin(double *pos, const double *mom, const double *F, double *Fold, const double *mass, double dt,
                     const double *posb, double4 *wpos, const double4 *wposb, int nat, unsigned long long *maxd2,
                     const int *step, hipStream_t s);
void launch_md_end(double *mom, const double *F, const double *Fold, double dt, double *slot, const double *pos,
                   int nat, const int *step, hipStream_t s);
void launch_md_load(const double *slot, double *pos, double *mom, double *F, int nat, hipStream_t s);
void launch_md_ekin(const double *mom, const double *mass, const int *off, int nsys, double *ekin, hipStream_t s);

}  // namespace hdnnp
