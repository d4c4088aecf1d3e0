// Molecular dynamics of many systems on the device: velocity Verlet
// (hdnnpExample.f90:58-72) with the forces of the engine, positions and
// momenta resident in device memory, and a Verlet skin: neighbour lists
// within rc + skin are reused while no atom has moved more than skin / 2
// since they were built (skin = 0 builds them at every step, as the
// reference does). Systems and evaluations may share one engine: a system
// whose lists another call has replaced since its build builds them again.
//
// Steps that reuse the lists run in chunks without waiting for the device:
// each step's largest displacement is recorded on the device and its state
// (positions, momenta, forces) saved, and the chunk is checked once at its
// end. From the first step that moved an atom beyond skin/2 on, the chunk is
// undone (the state after the step before it restored) and stepping goes on
// one step at a time, checked before each force evaluation, up to the list
// build. The trajectory is the one of checking every step. Chunks are sized
// from the steps between the last two builds.
//
// Within a chunk every step launches the same kernels on the same buffers
// (a step counter on the device selects the displacement and state slots),
// so after the chunk's first step the rest are replays of one captured HIP
// graph of a step. The graph is captured again when the engine's launch
// arguments change (a list build that moved or resized buffers). Setting
// HDNNP_MD_GRAPHS=0 in the environment launches every step directly.
#pragma once

#include <vector>

#include "engine.hpp"

namespace hdnnp {

class MdSystems {
  public:
    // nsys systems of nat[k] atoms; Z, pos, mom, mass over all atoms
    // (pos/mom [sum nat][3], mass per atom in electron masses); periodic[k],
    // lat[9k..] as for Engine::evaluate_batch_host; dt in atomic time units.
    MdSystems(Engine &eng, int nsys, const int *nat, const int *Z, const double *pos, const double *mom,
              const double *mass, const int *periodic, const double *lat, double dt, double skin);
    ~MdSystems();
    MdSystems(const MdSystems &) = delete;
    MdSystems &operator=(const MdSystems &) = delete;

    // nsteps velocity-Verlet steps; epot/ekin [nsys] of the last step (may be null)
    void step(int nsteps, double *epot, double *ekin);
    void get(double *pos, double *mom, double *force);   // host copies (any may be null)
    long long steps() const { return steps_; }
    long long builds() const { return builds_; }
    int nsys() const { return (int)st_.size(); }
    size_t device_bytes() const;   // the systems' own device state (not the engine's)
    int nat() const { return nat_; }

    long long rollbacks() const { return rollbacks_; }
    long long graph_captures() const { return captures_; }
    long long graph_replays() const { return replays_; }

  private:
    static constexpr int kMaxChunk = 64;
    void forces(double *epot);
    void init(int nsys, const int *nat, const int *Z, const double *pos, const double *mom, const double *mass,
              const int *periodic, const double *lat);
    void release();               // every device and pinned buffer, the graph
    void careful_step();          // one step, the skin checked before its forces
    int chunk(int k);             // up to k steps on reused lists; the steps kept
    void step_launch(bool timed); // one chunk step's kernels (the captured sequence)
    bool replay(int n);           // n more chunk steps by the captured graph; false: not available

    Engine &eng_;
    std::vector<StructIn> st_;
    int nat_ = 0;
    double dt_, skin_;
    DevMem Z_, pos_, mom_, mass_, F_, Fold_, posb_, wposb_, off_, ekin_, maxd2_, snap_, d2_;
    unsigned long long *d2_host_ = nullptr;   // pinned, kMaxChunk entries
    int maxk_ = kMaxChunk;         // chunk limit (the saved states stay within 256 MiB)
    int since_ = 0, interval_ = 0;  // steps since the last build, and between the two before
    bool careful_ = true;          // step one at a time until the next list build
    long long rollbacks_ = 0;
    DevMem ctr_;                   // step counter of the running chunk
    bool graphs_ = true;
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t exec_ = nullptr;
    std::vector<char> key_;        // engine launch key the graph was captured with
    long long captures_ = 0, replays_ = 0;
    std::vector<double> epot_, ekin_h_;
    bool built_ = false;
    unsigned long long gen_ = 0;   // engine build generation of this system's lists
    long long steps_ = 0, builds_ = 0;
};

}  // namespace hdnnp
