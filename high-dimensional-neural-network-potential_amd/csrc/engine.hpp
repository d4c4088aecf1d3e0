// Evaluation engine: owns the device copy of one potential and the device
// work buffers, and runs hdnnpEnergyAndForces on the GPU.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <shared_mutex>
#include <string>
#include <vector>

#include "device.hpp"
#include "runner.hpp"

namespace hdnnp {

// HIP graph capture (MdSystems::replay) and the library's calls that a capture
// in another thread makes illegal, or that invalidate it: allocations and
// frees, synchronous copies on the legacy stream, stream, event and graph
// creation and destruction. A capture holds the mutex exclusively; those calls
// hold it shared through HipUnsafe. The thread that captures takes no shared
// lock (it already holds the mutex); nested HipUnsafe scopes lock once.
std::shared_mutex &capture_mutex();
class HipUnsafe {
  public:
    HipUnsafe();
    ~HipUnsafe();
    HipUnsafe(const HipUnsafe &) = delete;
    HipUnsafe &operator=(const HipUnsafe &) = delete;
    // the calling thread holds the mutex exclusively for a capture
    struct Capture {
        Capture();
        ~Capture();
        std::unique_lock<std::shared_mutex> lk;
    };

  private:
    bool locked_ = false, counted_ = false;
};

// Device scratch of the entry point that allocates nothing
// (Engine::evaluate_batch_device_noalloc): bytes per device.
constexpr size_t kScratchBytes = size_t(128) << 20;

// grow-only device allocation; inside evaluate_batch_device_noalloc it carves
// from the device scratch instead (no allocation, nothing freed)
struct DevMem {
    void *p = nullptr;
    size_t bytes = 0;
    void ensure(size_t b);
    void release();
    template <typename T> T *as() const { return (T *)p; }
};

// One structure of an evaluation as given by the caller.
struct StructIn {
    int nat = 0;
    int periodic = 0;
    double lat[9] = {};    // periodic: lattice, lat[3k+a] = component a of vector k
    double bbox[6] = {};   // nonperiodic: min and max of the raw positions
};

class Engine {
  public:
    // members[0] is the potential; further members form a committee with it
    // (their networks on the first potential's symmetry functions)
    Engine(std::vector<HostModel> members, int precision);
    int committee_size() const { return 1 + (int)members_.size(); }
    // last evaluation: member energies [nstruct][N] and their standard deviation
    const std::vector<double> &member_energies() const { return member_E_; }
    const std::vector<double> &energy_std() const { return std_E_; }
    ~Engine();
    Engine(const Engine &) = delete;
    Engine &operator=(const Engine &) = delete;

    const HostModel &model() const { return model_; }
    int precision() const { return precision_; }

    // host arrays in and out (hdnnp_evaluate)
    void evaluate_host(int nat, const int *Z, const double *pos, int periodic, const double *lat,
                       double *energy, double *force, double *dEdLat);
    // nstruct structures concatenated: nat[k] atoms each, Z/pos/force host
    // arrays over all atoms, periodic[k], lat[9k..] (periodic ones),
    // energy[k], dEdLat[9k..] (may be null)
    void evaluate_batch_host(int nstruct, const int *nat, const int *Z, const double *pos, const int *periodic,
                             const double *lat, double *energy, double *force, double *dEdLat);
    // the same with Z, pos and force in device memory (nat, periodic, lat,
    // energy, dEdLat on the host)
    void evaluate_batch_device(int nstruct, const int *nat, const int *Z_dev, const double *pos_dev,
                               const int *periodic, const double *lat, double *energy, double *force_dev,
                               double *dEdLat, hipStream_t stream);
    // evaluate_batch_device with every work buffer carved from the device
    // scratch (kScratchBytes per device): no device or pinned allocation, no
    // memset. Needs the context's device state (set up by create when a GPU is
    // visible); inputs that need more scratch are refused (capacity error).
    // Calls on one device are serialised. The context's own work buffers are
    // left as they were, and stage outputs of the call are not kept.
    void evaluate_batch_device_noalloc(int nstruct, const int *nat, const int *Z_dev, const double *pos_dev,
                                       const int *periodic, const double *lat, double *energy, double *force_dev,
                                       double *dEdLat, hipStream_t stream);
    // device arrays Z, pos, force; lat/energy/dEdLat on the host
    void evaluate_device(int nat, const int *Z_dev, const double *pos_dev, int periodic,
                         const double *lat, double *energy, double *force_dev, double *dEdLat,
                         hipStream_t stream);

    // ---- the reference's stage API (hdnnpCalcNeighborLists,
    // hdnnpCalcSymfunctions, hdnnpCalcNNs; hdnnps.f90:373-562) ----
    // The engine keeps one stage: the lists of one structure (host arrays in).
    // stage_build builds them and returns the stage's id; stage_current tells
    // whether they are still the engine's (no other build since).
    int64_t stage_build(int nat, const int *Z, const double *pos, int periodic, const double *lat);
    bool stage_current(int64_t id) const;
    // per atom nneis[i] = 1 + its entries (slot 1 = the atom itself); with
    // ineis / neis (may be null) the lists in the reference's layout and
    // order: 1-based atom numbers, positions of the images (neighborLists.f90:
    // 118-127, 211-213)
    void stage_neighbors(int32_t *nneis, int32_t *ineis, double *neis);
    // scaled G[i * ld + s] of the current stage: together != 0 through the
    // fused values pass (calcAllSymmetryFunctions), else one function at a
    // time (symmetryFunction2 / 3); dsf (may be null) gets the scaled dG/dR
    // rows of every atom, function and slot (both ways through the per-function
    // kernel)
    void stage_symfunctions(int together, int ld, double *G, double *dsf);
    // networks of `net` (one potential, same elements and SF counts, same
    // precision) on the G given (scaled, G[i * ld + s]) and this engine's
    // current stage: E, forces, dEdLat (may be null)
    void stage_nns(Engine &net, int ld, const double *G, double *energy, double *force, double *dEdLat);

    int64_t neighbor_count();
    void get_neighbors(int64_t cap, int32_t *ent);
    void get_symfunctions(int ld, double *G, double *dEdG, double *Ei);
    void timings(double *ms6) const;

    // With graphs on, evaluations of up to kSmallSortAtoms atoms on the
    // context's own stream (build and finish: a dozen small kernels) are
    // replayed from captured graphs when the same launch repeats: the second
    // evaluation with the launch parameters and buffers of the one before it
    // is captured, and later ones with any of the last kEvalGraphs captured
    // launches replay it. Off by default (a replay measured slower than the
    // direct launches, whose host side overlaps the kernels); on with
    // HDNNP_EVAL_GRAPHS=1 or set_eval_graphs. Counts of captures and replays.
    void set_eval_graphs(bool on);
    bool eval_graphs() const { return eval_graphs_on_; }
    void eval_graph_counts(int64_t *captures, int64_t *replays) const;

    // Stage timings (timings(): HIP events between the stages). Off by
    // default: the seven events of an evaluation are packets of their own on
    // the stream and cost a small evaluation ~10-20 us (8-atom molecule 116
    // against 104 us, 6^3 cell 0.207 against 0.184 ms). On with
    // HDNNP_STAGE_TIMINGS=1 or set_stage_timings; timings() holds zeros for
    // an evaluation launched without them.
    void set_stage_timings(bool on) { stage_timing_ = on; }
    bool stage_timings() const { return stage_timing_; }

  private:
    static constexpr int kEvalGraphs = 4;
    struct EvalGraph {
        std::vector<char> key;
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        unsigned long long used = 0;
    };
    std::vector<EvalGraph> eval_graphs_;
    std::vector<char> eval_seen_;        // launch of the last direct evaluation
    bool eval_graphs_on_ = false;
    bool stage_timing_ = false;
    bool ev_recorded_ = false;          // the last launched evaluation recorded ev_
    unsigned long long eval_tick_ = 0;
    int64_t eval_captures_ = 0, eval_replays_ = 0;
    // the prepared evaluation's kernels: replayed, captured or launched
    void launch_evaluation(double *force_dev, hipStream_t s);
    std::vector<char> eval_key() const;
    EvalGraph *capture_evaluation(double *force_dev, hipStream_t s, std::vector<char> key);
    void drop_eval_graphs();
    void init_device();
    void d2h(void *dst, const void *src, size_t bytes);
    // The structures st (atoms concatenated in Z_dev / pos_dev / force_dev).
    // retry: how often an evaluation may be repeated with a larger neighbour capacity
    void run(std::vector<StructIn> &st, const int *Z_dev, const double *pos_dev, double *force_dev,
             hipStream_t s, double *energy, double *dEdLat, int retry);
  public:
    // The two phases of run. build: cell grids, sorts and neighbour lists
    // within rc_list (>= the potential's cutoff) of the positions in pos_dev
    // (wrapped copies in wpos()). finish: symmetry functions of the current
    // wrapped positions, networks, forces and sums; false (nothing written)
    // when an atom had more neighbours than the list capacity -- build again.
    // launch = false: the host part only (grids, tables and their upload,
    // buffer sizes); build_launch then queues the kernels (timed = false:
    // no timing events, the form a captured graph holds)
    void build(const std::vector<StructIn> &st, const int *Z_dev, const double *pos_dev, hipStream_t s,
               double rc_list, bool launch = true);
    void build_launch(hipStream_t s, bool timed);
    bool finish(double *force_dev, hipStream_t s, double *energy, double *dEdLat);
    // finish in two parts: the kernels (queued on s, nothing waited for),
    // then the read-back, checks and sums (waits for s)
    // timed = false: no timing events (the form a captured graph can hold)
    void finish_launch(double *force_dev, hipStream_t s, bool timed = true);
    // Everything finish_launch's kernels take as arguments (launch
    // parameters and buffer addresses): a graph captured from finish_launch
    // is valid while this stays the same.
    std::vector<char> launch_key() const;
    bool finish_collect(hipStream_t s, double *energy, double *dEdLat);
    double4 *wpos() { return wpos_.as<double4>(); }
    // bounding boxes of the nonperiodic structures (for build), from device positions
    void device_bboxes(std::vector<StructIn> &st, const double *pos_dev, hipStream_t s);
    int prepared_nat() const { return prep_.nat; }
    bool prepared() const { return prep_.valid; }   // a built list that finish can use
    // count of builds so far: a caller that reuses lists checks that no other
    // evaluation on this engine has built since its own build
    unsigned long long build_generation() const { return build_gen_; }
    void ensure_device() { if (!dev_ready_) init_device(); }
    hipStream_t stream() const { return stream_; }

    // Tile of the symmetry-function kernels: atoms per wavefront, 1 (the
    // default), 2 or 4. It is one per context, so every evaluation on the
    // context uses it; a staging capacity or LDS size that does not take it
    // falls back to a smaller one.
    void set_tile(int atoms_per_wave);
    int tile() const { return tile_; }
    // Times the tiles on a sample (nstruct structures, host arrays as for
    // evaluate_batch_host): the two symmetry-function passes by HIP events,
    // best of 3 after a warm-up. Keeps the fastest as the context's tile and
    // returns it; ms[3] (may be null) gets the time of tiles 1, 2, 4 (0 for
    // one the sample's staging capacity does not take).
    int autotune(int nstruct, const int *nat, const int *Z, const double *pos, const int *periodic,
                 const double *lat, double *ms);

    // Device memory of the work buffers: bytes held, and their release (the
    // next evaluation allocates again; the device copy of the potential stays)
    size_t device_bytes() const;
    void release_work();
    // Device bytes an evaluation needs, from the sizes build and compute give
    // their buffers: structures are added one at a time (lists within
    // rc_list; a nonperiodic structure needs its bounding box)
    struct Footprint {
        long nat = 0, ncell = 0, ps = 0, nstruct = 0;
        double est = 0;   // largest neighbour-count estimate
    };
    void footprint_add(Footprint &f, const StructIn &in, double rc_list) const;
    // slack: the buffers' growth room and the sort temporaries on top
    size_t footprint_bytes(const Footprint &f, bool slack = true) const;

  private:
    void *staging(size_t bytes);
    void *io_staging(size_t bytes);
    std::vector<DevMem *> work_buffers();
    template <typename R>
    void compute(Work &w, const Geo &g, hipStream_t s, bool timed);

    HostModel model_;
    std::vector<HostModel> members_;          // committee members 2..N
    std::vector<DevModel> dmm_;               // device models of all members (dmm_[0] = dm_)
    std::vector<int> maxw_m_;                 // widest layer of each member's networks
    std::vector<ElemInfo> ei_host_;
    std::vector<double> member_E_, std_E_;
    int precision_;
    bool dev_ready_ = false;
    hipStream_t stream_ = nullptr;
    hipEvent_t ev_[7] = {};
    double ms_[6] = {0, 0, 0, 0, 0, 0};
    DevModel dm_{};
    std::vector<DevMem> model_mem_;
    bool elem_valid_[kMaxEl] = {};
    std::string elem_why_[kMaxEl];
    int mlp_maxw_ = 0, mlp_sumc_ = 0;

    // work buffers
    DevMem posin_, fout_;   // host-array evaluations: positions, then Z (256-aligned); forces
    DevMem eidx_, wpos_, cellkey_, cellkey_sorted_, cell_atoms_, iota_, cell_start_, cell_end_;
    DevMem count_, perm_, Ei_, vir_, misc_;
    DevMem ent_j_, ent_shift_, ent_s_, rev_, pos_start_, gent_, G_, dEdG_, cub_tmp_, elem_chunks_;
    DevMem trip_, tcount_;   // triplets pass A hands to pass B
    DevMem geo_, sid_, csid_, chunks_, sc_, res_, partial_, bbox_;
    std::vector<StructGeo> hgeo_;
    struct Prepared {
        Work w{};
        Geo g{};
        std::vector<StructIn> st;
        int nat = 0, nchunk = 0, lds_lim = 0, ncell = 0, bits = 1;
        size_t o_res = 0, o_flags = 0, res_b = 0;
        bool small_sort = false;
        bool count_sort = false;   // cells by counting sort (else radix)
        bool valid = false;
    } prep_;
    unsigned long long build_gen_ = 0;
    int tile_ = 1;
    std::vector<char> uploaded_;   // the staged tables of the last upload
    void *pinned_ = nullptr;
    size_t pinned_bytes_ = 0;
    void *stage_ptr_ = nullptr;          // what staging() returned last
    std::vector<char> host_stage_;       // staging without pinned memory (no-allocation calls)
    int dev_ = -1;                       // the device of init_device
    void *scratch_ = nullptr;            // that device's scratch (hipGetSymbolAddress)
    // pinned staging of a small host-array evaluation's inputs and forces
    // (apart from pinned_, which build() rewrites while those copies run)
    void *io_pinned_ = nullptr;
    size_t io_pinned_bytes_ = 0;
    // a force read-back that finish_collect enqueues before its one
    // synchronisation (host-array evaluations)
    struct {
        double *host = nullptr;
        const double *dev = nullptr;
        size_t bytes = 0;
    } pending_force_;
    // stage API state
    int64_t stage_id_ = 0, stage_next_ = 0;
    unsigned long long stage_gen_ = 0;
    StructIn stage_in_{};
    std::vector<double> stage_pos_;     // the stage's input positions
    bool skip_values_ = false;          // compute: G is already in G_ (stage_nns)
    bool trip_valid_ = false;           // trip_ holds the values pass's triplets of the current build
    bool tblk_valid_ = false;           // ... and their block tables (the FP32 values pass in block order)
    DevMem sfpar_, sfpar_off_, stage_G_, stage_dsf_, stage_off_;   // per-function path
    int last_nat_ = 0, last_cap_ = 0, last_lim_ = 0;
    int64_t last_total_ = 0;   // -1: not yet summed
    int cap_hint_ = 0;         // most neighbours of one atom seen so far
};

// host copies of reference 3x3 routines (exact operation order)
void inv3(const double A[9], double B[9]);
double det3(const double A[9]);

}  // namespace hdnnp
