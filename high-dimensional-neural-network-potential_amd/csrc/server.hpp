// Request server: continuous batching of energy/force/stress evaluations and
// long-lived molecular-dynamics sessions on one GPU.
//
// Lanes are worker threads, each with an engine of its own (its own HIP
// stream and work buffers). A free lane takes every queued evaluation that
// fits (in arrival order, up to max_batch_atoms atoms and within the device
// memory budget) and evaluates them together as one batch; when none waits,
// it takes the oldest request of an MD session. An MD request that has waited
// longer than kMdAging goes first while no other lane runs one, so under
// evaluation load the sessions keep one lane and never starve, and the
// evaluations keep the others. Requests that arrive while the lanes are
// busy wait for the next free lane; lanes run at the same time on their own
// streams, so their work overlaps on the device.
//
// Admission by device memory: the work buffers each lane holds (or the
// estimate of the batch it runs, when larger) and the device state of every
// MD session count against the budget. Work is started only when it fits;
// idle lanes give their buffers back when waiting work does not fit
// otherwise, and evaluations queued behind work that waits for memory wait
// too (no overtaking by smaller requests). A request or session that needs
// more than the whole budget is refused with a capacity error.
//
// An MD session is an MdSystems of one system with an engine of its own, so
// its positions, momenta and neighbour lists (with a Verlet skin) stay on the
// device between requests. Its requests run one at a time and in order, on
// whichever lane takes them.
//
// Results do not depend on the mix: a structure in a batch gets the energy,
// forces and dE/dh it gets alone (batched evaluation equals the single one
// bit for bit), and a session steps as it would alone.
#pragma once

#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "engine.hpp"
#include "md.hpp"

namespace hdnnp {

class Server {
  public:
    // budget_bytes 0: 80% of the device memory free at creation;
    // max_batch_atoms <= 0: 262144
    Server(std::vector<HostModel> members, int precision, int nlanes, size_t budget_bytes, long max_batch_atoms);
    ~Server();
    Server(const Server &) = delete;
    Server &operator=(const Server &) = delete;

    // queue an evaluation (inputs copied); returns its ticket
    long long submit(int nat, const int *Z, const double *pos, int periodic, const double *lat);
    // wait for an evaluation (timeout_s <= 0: no limit); false on time-out
    // (the ticket stays valid); throws the request's error. energy_std: the
    // committee's standard deviation of the energy (0 for one potential)
    bool wait(long long ticket, double timeout_s, double *energy, double *force, double *dEdLat, double *energy_std,
              double *latency_ms);
    // cancel a queued evaluation or MD-step request: true when it had not
    // started (its wait then throws kCancelled), false when it runs or is done
    bool cancel(long long ticket);
    static constexpr int kCancelled = 8;

    // open an MD session (opened on a lane once it fits the budget); its id
    long long md_open(int nat, const int *Z, const double *pos, const double *mom, const double *mass, int periodic,
                      const double *lat, double dt, double skin);
    // queue nsteps velocity-Verlet steps of a session (0: none, only the
    // energies of its current state); returns the ticket
    long long md_step(long long session, int nsteps);
    bool md_wait(long long ticket, double timeout_s, double *epot, double *ekin, double *latency_ms);
    // state of a session after all its queued requests
    void md_get(long long session, double *pos, double *mom, double *force);
    // close a session once its queued requests are done; frees its device state
    void md_close(long long session);

    // [0] evaluations done, [1] their atoms, [2] MD requests done, [3] MD
    // atom-steps, [4]/[5] p50/p99 evaluation latency (ms), [6]/[7] p50/p99
    // MD request latency (ms), [8] batches, [9] mean structures per batch,
    // [10] peak reserved device bytes, [11] budget, [12] seconds since the
    // last reset, [13] most lanes busy at once, [14] requests refused, [15]
    // MD steps
    static constexpr int kStats = 16;
    void stats(double *out, bool reset);
    // an MD request waiting longer than this goes before evaluations (while
    // no other lane runs an MD request)
    static constexpr std::chrono::milliseconds kMdAging{100};

  private:
    using Clock = std::chrono::steady_clock;
    struct Session;
    struct Job {
        enum Kind { Eval, Open, Step } kind = Eval;
        long long ticket = 0;
        StructIn st;
        std::vector<int> Z;
        std::vector<double> pos, mom, mass;
        double dt = 0, skin = 0;
        std::shared_ptr<Session> sess;
        int nsteps = 0;
        size_t est = 0;             // estimated device bytes (Open)
        double energy = 0, estd = 0, dEdLat[9] = {}, epot = 0, ekin = 0;
        std::vector<double> force;
        bool done = false;
        int code = 0, fcode = 0;    // fcode/ferr: refused while queued
        std::string err, ferr;
        Clock::time_point t_submit, t_done;
    };
    struct Session {
        long long id = 0;
        int nat = 0;
        std::unique_ptr<Engine> eng;
        std::unique_ptr<MdSystems> md;
        bool busy = false, opened = false, failed = false, closing = false;
        int queued = 0;             // requests queued or running
        size_t bytes = 0;           // reserved device bytes
        int code = 0;
        std::string err;
    };
    struct Lane {
        std::unique_ptr<Engine> eng;
        std::thread th;
        size_t held = 0, reserved = 0;
        bool running = false, trim = false;
    };

    void lane_main(int l);
    bool pick(int l, std::vector<std::shared_ptr<Job>> &batch, std::shared_ptr<Job> &one);
    void run_batch(int l, std::vector<std::shared_ptr<Job>> &batch);
    void run_md(int l, const std::shared_ptr<Job> &job);
    void finish_job(Job &j, int code, const std::string &err);   // with mu_ held
    size_t reserved_total() const;
    void request_trims(int l, size_t deficit);
    std::shared_ptr<Job> take(long long ticket, double timeout_s, bool md);
    std::shared_ptr<Session> session(long long id);

    std::vector<HostModel> models_;
    int precision_;
    int dev_ = 0;
    size_t budget_ = 0;
    long max_batch_atoms_ = 0;
    double rc_ = 0;
    std::unique_ptr<Engine> sizing_;     // footprint estimates at submission (never touches the device)
    std::vector<std::unique_ptr<Lane>> lanes_;

    std::mutex mu_;
    std::condition_variable cv_work_, cv_done_;
    std::deque<std::shared_ptr<Job>> queue_;
    std::unordered_map<long long, std::shared_ptr<Job>> jobs_;
    std::unordered_map<long long, std::shared_ptr<Session>> sessions_;
    long long next_ticket_ = 1, next_session_ = 1;
    bool stop_ = false;
    size_t sessions_bytes_ = 0, peak_ = 0;
    std::vector<double> lat_eval_, lat_md_;
    double atoms_eval_ = 0, atom_steps_ = 0, md_steps_ = 0, batches_ = 0, batch_structs_ = 0, refused_ = 0;
    int running_ = 0, max_running_ = 0;
    int md_running_ = 0;                 // lanes running an MD request
    Clock::time_point t0_;
};

}  // namespace hdnnp
