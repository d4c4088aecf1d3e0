// Request server (see server.hpp): lanes, the queue, admission by device
// memory, MD sessions and latency statistics.
#include "server.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace hdnnp {

#define SVCHK(x)                                                                               \
    do {                                                                                       \
        hipError_t e_ = (x);                                                                   \
        if (e_ != hipSuccess)                                                                  \
            throw Error(5, std::string("HIP error ") + hipGetErrorString(e_) + " at " #x);     \
    } while (0)

static constexpr size_t kMiB = size_t(1) << 20;

// the structure of a request: lattice (periodic) or bounding box of the
// positions (nonperiodic, neighborLists.f90:131-138)
static StructIn structure_of(int nat, const double *pos, int periodic, const double *lat)
{
    StructIn in;
    in.nat = nat;
    in.periodic = periodic ? 1 : 0;
    if (in.periodic) {
        if (!lat) throw Error(1, "periodic structure without lattice");
        std::memcpy(in.lat, lat, sizeof in.lat);
    } else if (nat > 0) {
        for (int a = 0; a < 3; ++a) in.bbox[a] = in.bbox[3 + a] = pos[a];
        for (int i = 1; i < nat; ++i)
            for (int a = 0; a < 3; ++a) {
                in.bbox[a] = std::min(in.bbox[a], pos[3 * i + a]);
                in.bbox[3 + a] = std::max(in.bbox[3 + a], pos[3 * i + a]);
            }
    }
    for (double v : in.bbox)
        if (!std::isfinite(v)) throw Error(1, "positions must be finite");
    return in;
}

Server::Server(std::vector<HostModel> members, int precision, int nlanes, size_t budget_bytes, long max_batch_atoms)
    : models_(std::move(members)), precision_(precision)
{
    if (nlanes < 1 || nlanes > 64) throw Error(1, "lanes must be 1..64");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) throw Error(5, "no HIP device available");
    SVCHK(hipGetDevice(&dev_));
    sizing_ = std::make_unique<Engine>(models_, precision_);   // validates the potential(s)
    rc_ = sizing_->model().maxrc;
    if (!(rc_ > 0)) throw Error(3, "potential has no symmetry functions (cutoff 0)");
    if (budget_bytes == 0) {
        size_t fr = 0, tot = 0;
        SVCHK(hipMemGetInfo(&fr, &tot));
        budget_bytes = fr / 10 * 8;
    }
    budget_ = budget_bytes;
    max_batch_atoms_ = max_batch_atoms > 0 ? max_batch_atoms : 262144;
    for (int l = 0; l < nlanes; ++l) {
        lanes_.push_back(std::make_unique<Lane>());
        lanes_.back()->eng = std::make_unique<Engine>(models_, precision_);
        lanes_.back()->eng->ensure_device();
    }
    t0_ = Clock::now();
    try {
        for (int l = 0; l < nlanes; ++l) lanes_[l]->th = std::thread(&Server::lane_main, this, l);
    } catch (...) {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_work_.notify_all();
        for (auto &L : lanes_)
            if (L->th.joinable()) L->th.join();
        throw;
    }
}

Server::~Server()
{
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
    }
    cv_work_.notify_all();
    for (auto &L : lanes_)
        if (L->th.joinable()) L->th.join();
    {
        std::lock_guard<std::mutex> lk(mu_);
        for (auto &jp : queue_) finish_job(*jp, 5, "server closed");
        queue_.clear();
        jobs_.clear();
    }
    (void)hipSetDevice(dev_);
    sessions_.clear();   // each session's MD state, then its engine
    lanes_.clear();
}

size_t Server::reserved_total() const
{
    size_t b = sessions_bytes_;
    for (const auto &L : lanes_) b += L->reserved;
    return b;
}

// ask idle lanes other than l to give back their buffers until about
// deficit bytes are free
void Server::request_trims(int l, size_t deficit)
{
    size_t freed = 0;
    for (int k = 0; k < (int)lanes_.size() && freed < deficit; ++k) {
        Lane &K = *lanes_[k];
        if (k == l || K.running || K.trim || K.held == 0) continue;
        K.trim = true;
        freed += K.held;
    }
    if (freed > 0) cv_work_.notify_all();
}

void Server::finish_job(Job &j, int code, const std::string &err)
{
    j.code = code;
    j.err = err;
    j.done = true;
    j.t_done = Clock::now();
    const double ms = std::chrono::duration<double, std::milli>(j.t_done - j.t_submit).count();
    if (j.kind == Job::Eval) {
        if (!code) {
            lat_eval_.push_back(ms);
            atoms_eval_ += j.st.nat;
        }
    } else {
        if (j.kind == Job::Step && !code) {
            lat_md_.push_back(ms);
            atom_steps_ += (double)j.sess->nat * j.nsteps;
            md_steps_ += j.nsteps;
        }
        j.sess->queued--;
    }
    std::vector<int>().swap(j.Z);
    std::vector<double>().swap(j.pos);
    std::vector<double>().swap(j.mom);
    std::vector<double>().swap(j.mass);
    cv_done_.notify_all();
}

// The next work of a lane. Evaluations (latency-bound requests) go first:
// every queued one that fits, in arrival order, as one batch. MD requests run
// when no evaluation waits, oldest first and each session's in order; one
// that has waited longer than kMdAging goes before the evaluations while no
// other lane runs an MD request, so that sessions never starve and, when they
// alone would fill every lane, evaluations still keep all lanes but one.
// Called with mu_ held by an idle lane.
bool Server::pick(int l, std::vector<std::shared_ptr<Job>> &batch, std::shared_ptr<Job> &one)
{
    Lane &L = *lanes_[l];
    const size_t others = reserved_total() - L.reserved;
    std::vector<Job *> failed;

    // the oldest runnable MD request
    std::shared_ptr<Job> md;
    {
        std::vector<const Session *> seen;
        bool open_blocked = false;
        for (auto &jp : queue_) {
            Job &j = *jp;
            if (j.kind == Job::Eval) continue;
            const Session *S = j.sess.get();
            if (std::find(seen.begin(), seen.end(), S) != seen.end()) continue;   // only its oldest request
            seen.push_back(S);
            if (j.kind == Job::Step) {
                if (S->failed) {
                    j.fcode = S->code;
                    j.ferr = S->err;
                    failed.push_back(&j);
                    continue;
                }
                if (!S->opened || S->busy) continue;
                md = jp;
                break;
            }
            if (open_blocked) continue;                 // sessions open in arrival order
            if (reserved_total() + j.est > budget_) {
                request_trims(l, reserved_total() + j.est - budget_);
                if (L.held > 0) L.trim = true;
                open_blocked = true;
                continue;
            }
            md = jp;
            break;
        }
    }
    const bool md_overdue = md && md_running_ == 0 && Clock::now() - md->t_submit > kMdAging;

    // a batch of evaluations
    Engine::Footprint fp;
    long atoms = 0;
    size_t need_batch = 0;
    bool mem_blocked = false;
    for (auto &jp : queue_) {
        if (md_overdue) break;
        Job &j = *jp;
        if (j.kind != Job::Eval || mem_blocked) continue;
        if (!batch.empty() && atoms + j.st.nat > max_batch_atoms_) continue;
        Engine::Footprint f2 = fp;
        size_t est = 0;
        try {
            L.eng->footprint_add(f2, j.st, rc_);
            est = L.eng->footprint_bytes(f2);
        } catch (const Error &e) {
            j.fcode = e.code;
            j.ferr = e.what();
            failed.push_back(&j);
            continue;
        }
        const size_t need = std::max(L.held, est);
        if (others + need > budget_) {
            if (batch.empty()) {
                if (est > budget_) {         // does not fit even on an empty server
                    j.fcode = 6;
                    j.ferr = "the request needs about " + std::to_string(est / kMiB) +
                             " MiB of device memory; the server's budget is " + std::to_string(budget_ / kMiB) +
                             " MiB";
                    refused_ += 1;
                    failed.push_back(&j);
                    continue;
                }
                // idle lanes give back their buffers, then this lane its own excess
                request_trims(l, others + need - budget_);
                if (L.held > est && others + est <= budget_) L.trim = true;
            }
            mem_blocked = true;          // later evaluations do not overtake it
            continue;
        }
        fp = f2;
        atoms += j.st.nat;
        need_batch = need;
        batch.push_back(jp);
    }
    if (!batch.empty()) {
        L.reserved = need_batch;
    } else if (md) {
        Job &j = *md;
        if (j.kind == Job::Open) {
            j.sess->bytes = j.est;
            sessions_bytes_ += j.est;
        }
        j.sess->busy = true;
        ++md_running_;
        one = md;
    }
    if (failed.empty() && batch.empty() && !one) return false;
    std::vector<const Job *> gone(failed.begin(), failed.end());
    for (auto &jp : batch) gone.push_back(jp.get());
    if (one) gone.push_back(one.get());
    for (Job *j : failed) finish_job(*j, j->fcode, j->ferr);
    queue_.erase(std::remove_if(queue_.begin(), queue_.end(),
                                [&](const std::shared_ptr<Job> &p) {
                                    return std::find(gone.begin(), gone.end(), p.get()) != gone.end();
                                }),
                 queue_.end());
    return !batch.empty() || one;
}

void Server::lane_main(int l)
{
    Lane &L = *lanes_[l];
    (void)hipSetDevice(dev_);
    for (;;) {
        std::vector<std::shared_ptr<Job>> batch;
        std::shared_ptr<Job> one;
        bool trim = false;
        {
            std::unique_lock<std::mutex> lk(mu_);
            for (;;) {
                if (stop_) return;
                if (L.trim) {
                    trim = true;
                    break;
                }
                if (pick(l, batch, one)) break;
                if (L.trim) {
                    trim = true;
                    break;
                }
                cv_work_.wait(lk);
            }
            if (!trim) {
                L.running = true;
                max_running_ = std::max(max_running_, ++running_);
                peak_ = std::max(peak_, reserved_total());
            }
        }
        if (trim) {
            try {
                L.eng->release_work();
            } catch (const std::exception &) {
            }
            std::lock_guard<std::mutex> lk(mu_);
            L.trim = false;
            L.held = L.reserved = L.eng->device_bytes();
            cv_work_.notify_all();
            continue;
        }
        if (one)
            run_md(l, one);
        else
            run_batch(l, batch);
    }
}

void Server::run_batch(int l, std::vector<std::shared_ptr<Job>> &batch)
{
    Lane &L = *lanes_[l];
    Engine &eng = *L.eng;
    const int n = (int)batch.size();
    std::vector<int> nat(n), per(n), Z;
    std::vector<double> pos, lat(9 * (size_t)n, 0.0);
    std::vector<long> a0(n + 1, 0);
    for (int k = 0; k < n; ++k) {
        const Job &j = *batch[k];
        nat[k] = j.st.nat;
        per[k] = j.st.periodic;
        std::memcpy(&lat[9 * (size_t)k], j.st.lat, sizeof j.st.lat);
        Z.insert(Z.end(), j.Z.begin(), j.Z.end());
        pos.insert(pos.end(), j.pos.begin(), j.pos.end());
        a0[k + 1] = a0[k] + nat[k];
    }
    const long total = a0[n];
    std::vector<double> E(n, 0.0), F(3 * (size_t)std::max<long>(total, 1), 0.0), D(9 * (size_t)n, 0.0),
        sd(n, 0.0);
    std::vector<int> code(n, 0);
    std::vector<std::string> err(n);
    auto eval = [&](int k0, int cnt) {
        eng.evaluate_batch_host(cnt, &nat[k0], Z.data() + a0[k0], pos.data() + 3 * a0[k0], &per[k0], &lat[9 * k0],
                                &E[k0], F.data() + 3 * a0[k0], &D[9 * k0]);
        if (eng.committee_size() > 1)
            for (int k = 0; k < cnt; ++k) sd[k0 + k] = eng.energy_std()[k];
    };
    auto attempt = [&](int k0, int cnt) -> bool {
        try {
            eval(k0, cnt);
            return true;
        } catch (const Error &e) {
            if (cnt == 1) code[k0] = e.code, err[k0] = e.what();
        } catch (const std::bad_alloc &) {
            if (cnt == 1) code[k0] = 6, err[k0] = "host out of memory";
        } catch (const std::exception &e) {
            if (cnt == 1) code[k0] = 5, err[k0] = e.what();
        }
        return false;
    };
    // a failed batch is evaluated again one request at a time, so that an
    // error belongs to its own request only
    if (!attempt(0, n) && n > 1)
        for (int k = 0; k < n; ++k) attempt(k, 1);

    size_t held = 0;
    try {
        held = eng.device_bytes();
    } catch (...) {
    }
    std::lock_guard<std::mutex> lk(mu_);
    L.held = L.reserved = held;
    L.running = false;
    --running_;
    peak_ = std::max(peak_, reserved_total());
    batches_ += 1;
    batch_structs_ += n;
    for (int k = 0; k < n; ++k) {
        Job &j = *batch[k];
        if (!code[k]) {
            j.energy = E[k];
            j.estd = sd[k];
            std::memcpy(j.dEdLat, &D[9 * (size_t)k], sizeof j.dEdLat);
            j.force.assign(F.begin() + 3 * a0[k], F.begin() + 3 * a0[k + 1]);
        }
        finish_job(j, code[k], err[k]);
    }
    cv_work_.notify_all();
}

void Server::run_md(int l, const std::shared_ptr<Job> &jp)
{
    Job &j = *jp;
    Session &S = *j.sess;
    int code = 0;
    std::string err;
    double epot = 0, ekin = 0;
    try {
        if (j.kind == Job::Open) {
            S.eng = std::make_unique<Engine>(models_, precision_);
            const int nat = j.st.nat, per = j.st.periodic;
            S.md = std::make_unique<MdSystems>(*S.eng, 1, &nat, j.Z.data(), j.pos.data(), j.mom.data(), j.mass.data(),
                                               &per, per ? j.st.lat : nullptr, j.dt, j.skin);
            epot = 0;
        } else {
            S.md->step(j.nsteps, &epot, &ekin);
        }
    } catch (const Error &e) {
        code = e.code;
        err = e.what();
    } catch (const std::bad_alloc &) {
        code = 6;
        err = "host out of memory";
    } catch (const std::exception &e) {
        code = 5;
        err = e.what();
    }
    if (code && j.kind == Job::Open) {
        S.md.reset();
        S.eng.reset();
    }
    size_t held = 0;
    if (S.eng) held = S.eng->device_bytes() + (S.md ? S.md->device_bytes() : 0);
    std::lock_guard<std::mutex> lk(mu_);
    Lane &L = *lanes_[l];
    L.running = false;
    --running_;
    --md_running_;
    sessions_bytes_ -= S.bytes;
    S.bytes = held;
    sessions_bytes_ += held;
    peak_ = std::max(peak_, reserved_total());
    S.busy = false;
    if (code) {
        S.failed = true;
        S.code = code;
        S.err = err;
    } else if (j.kind == Job::Open) {
        S.opened = true;
    }
    j.epot = epot;
    j.ekin = ekin;
    finish_job(j, code, err);
    cv_work_.notify_all();
}

long long Server::submit(int nat, const int *Z, const double *pos, int periodic, const double *lat)
{
    if (nat < 0) throw Error(1, "nat < 0");
    if (nat > 0 && (!Z || !pos)) throw Error(1, "null array");
    auto j = std::make_shared<Job>();
    j->kind = Job::Eval;
    j->st = structure_of(nat, pos, periodic, lat);
    Engine::Footprint f;
    sizing_->footprint_add(f, j->st, rc_);   // checks the lattice
    if (sizing_->footprint_bytes(f) > budget_) {
        std::lock_guard<std::mutex> lk(mu_);
        refused_ += 1;
        throw Error(6, "the request needs about " + std::to_string(sizing_->footprint_bytes(f) / kMiB) +
                           " MiB of device memory; the server's budget is " + std::to_string(budget_ / kMiB) +
                           " MiB");
    }
    j->Z.assign(Z, Z + nat);
    j->pos.assign(pos, pos + 3 * (size_t)nat);
    j->t_submit = Clock::now();
    std::lock_guard<std::mutex> lk(mu_);
    if (stop_) throw Error(5, "server closed");
    j->ticket = next_ticket_++;
    jobs_[j->ticket] = j;
    queue_.push_back(j);
    cv_work_.notify_one();
    return j->ticket;
}

std::shared_ptr<Server::Job> Server::take(long long ticket, double timeout_s, bool md)
{
    std::unique_lock<std::mutex> lk(mu_);
    auto it = jobs_.find(ticket);
    if (it == jobs_.end()) throw Error(1, "unknown ticket " + std::to_string(ticket));
    std::shared_ptr<Job> jp = it->second;
    if ((jp->kind == Job::Step) != md)
        throw Error(1, md ? "not the ticket of an MD request" : "not the ticket of an evaluation");
    auto ready = [&] { return jp->done; };
    if (timeout_s > 0) {
        if (!cv_done_.wait_for(lk, std::chrono::duration<double>(timeout_s), ready)) return nullptr;
    } else {
        cv_done_.wait(lk, ready);
    }
    jobs_.erase(ticket);
    return jp;
}

bool Server::wait(long long ticket, double timeout_s, double *energy, double *force, double *dEdLat,
                  double *energy_std, double *latency_ms)
{
    std::shared_ptr<Job> j = take(ticket, timeout_s, false);
    if (!j) return false;
    if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(j->t_done - j->t_submit).count();
    if (j->code) throw Error(j->code, j->err);
    if (energy) *energy = j->energy;
    if (force && !j->force.empty()) std::memcpy(force, j->force.data(), sizeof(double) * j->force.size());
    if (dEdLat) std::memcpy(dEdLat, j->dEdLat, sizeof j->dEdLat);
    if (energy_std) *energy_std = j->estd;
    return true;
}

bool Server::cancel(long long ticket)
{
    std::lock_guard<std::mutex> lk(mu_);
    auto it = jobs_.find(ticket);
    if (it == jobs_.end()) throw Error(1, "unknown ticket " + std::to_string(ticket));
    const std::shared_ptr<Job> jp = it->second;
    if (jp->done) return false;
    auto q = std::find(queue_.begin(), queue_.end(), jp);
    if (q == queue_.end()) return false;   // a lane runs it
    queue_.erase(q);
    finish_job(*jp, kCancelled, "request cancelled");
    cv_work_.notify_all();                 // work queued behind it may fit now
    return true;
}

long long Server::md_open(int nat, const int *Z, const double *pos, const double *mom, const double *mass,
                          int periodic, const double *lat, double dt, double skin)
{
    if (nat < 1) throw Error(1, "an MD session needs at least one atom");
    if (!Z || !pos || !mom || !mass) throw Error(1, "null array");
    if (!(dt > 0) || !(skin >= 0)) throw Error(1, "dt must be > 0 and skin >= 0");
    for (int i = 0; i < nat; ++i)
        if (!(mass[i] > 0)) throw Error(1, "masses must be positive");
    auto j = std::make_shared<Job>();
    j->kind = Job::Open;
    j->st = structure_of(nat, pos, periodic, lat);
    Engine::Footprint f;
    sizing_->footprint_add(f, j->st, rc_ + skin);
    // the engine's buffers, the session's own state, the potential's tables
    j->est = sizing_->footprint_bytes(f) + (size_t)nat * 250 + 16 * kMiB;
    if (j->est > budget_) {
        std::lock_guard<std::mutex> lk(mu_);
        refused_ += 1;
        throw Error(6, "the session needs about " + std::to_string(j->est / kMiB) +
                           " MiB of device memory; the server's budget is " + std::to_string(budget_ / kMiB) +
                           " MiB");
    }
    j->Z.assign(Z, Z + nat);
    j->pos.assign(pos, pos + 3 * (size_t)nat);
    j->mom.assign(mom, mom + 3 * (size_t)nat);
    j->mass.assign(mass, mass + nat);
    j->dt = dt;
    j->skin = skin;
    j->t_submit = Clock::now();
    auto S = std::make_shared<Session>();
    S->nat = nat;
    j->sess = S;
    std::lock_guard<std::mutex> lk(mu_);
    if (stop_) throw Error(5, "server closed");
    S->id = next_session_++;
    S->queued = 1;
    j->ticket = next_ticket_++;
    sessions_[S->id] = S;
    queue_.push_back(j);
    cv_work_.notify_one();
    return S->id;
}

std::shared_ptr<Server::Session> Server::session(long long id)
{
    auto it = sessions_.find(id);
    if (it == sessions_.end()) throw Error(1, "unknown MD session " + std::to_string(id));
    return it->second;
}

long long Server::md_step(long long id, int nsteps)
{
    if (nsteps < 0) throw Error(1, "nsteps < 0");
    std::lock_guard<std::mutex> lk(mu_);
    if (stop_) throw Error(5, "server closed");
    std::shared_ptr<Session> S = session(id);
    if (S->closing) throw Error(1, "MD session " + std::to_string(id) + " is closing");
    if (S->failed) throw Error(S->code, S->err);
    auto j = std::make_shared<Job>();
    j->kind = Job::Step;
    j->sess = S;
    j->nsteps = nsteps;
    j->t_submit = Clock::now();
    j->ticket = next_ticket_++;
    S->queued++;
    jobs_[j->ticket] = j;
    queue_.push_back(j);
    cv_work_.notify_one();
    return j->ticket;
}

bool Server::md_wait(long long ticket, double timeout_s, double *epot, double *ekin, double *latency_ms)
{
    std::shared_ptr<Job> j = take(ticket, timeout_s, true);
    if (!j) return false;
    if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(j->t_done - j->t_submit).count();
    if (j->code) throw Error(j->code, j->err);
    if (epot) *epot = j->epot;
    if (ekin) *ekin = j->ekin;
    return true;
}

void Server::md_get(long long id, double *pos, double *mom, double *force)
{
    std::shared_ptr<Session> S;
    {
        std::unique_lock<std::mutex> lk(mu_);
        S = session(id);
        cv_done_.wait(lk, [&] { return S->queued == 0 && !S->busy; });
        if (S->failed) throw Error(S->code, S->err);
        S->busy = true;
    }
    int code = 0;
    std::string err;
    try {
        SVCHK(hipSetDevice(dev_));   // the session's device, from any client thread
        S->md->get(pos, mom, force);
    } catch (const Error &e) {
        code = e.code;
        err = e.what();
    }
    std::lock_guard<std::mutex> lk(mu_);
    S->busy = false;
    cv_work_.notify_all();
    cv_done_.notify_all();
    if (code) throw Error(code, err);
}

void Server::md_close(long long id)
{
    std::shared_ptr<Session> S;
    {
        std::unique_lock<std::mutex> lk(mu_);
        S = session(id);
        S->closing = true;
        cv_done_.wait(lk, [&] { return S->queued == 0 && !S->busy; });
        sessions_.erase(id);
        S->busy = true;
    }
    (void)hipSetDevice(dev_);
    S->md.reset();
    S->eng.reset();
    std::lock_guard<std::mutex> lk(mu_);
    sessions_bytes_ -= S->bytes;
    S->bytes = 0;
    cv_work_.notify_all();
}

// nearest-rank percentile
static double percentile(std::vector<double> v, double q)
{
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const size_t r = (size_t)std::ceil(q * (double)v.size());
    return v[std::min(v.size(), std::max<size_t>(r, 1)) - 1];
}

void Server::stats(double *out, bool reset)
{
    std::lock_guard<std::mutex> lk(mu_);
    const Clock::time_point now = Clock::now();
    out[0] = (double)lat_eval_.size();
    out[1] = atoms_eval_;
    out[2] = (double)lat_md_.size();
    out[3] = atom_steps_;
    out[4] = percentile(lat_eval_, 0.50);
    out[5] = percentile(lat_eval_, 0.99);
    out[6] = percentile(lat_md_, 0.50);
    out[7] = percentile(lat_md_, 0.99);
    out[8] = batches_;
    out[9] = batches_ > 0 ? batch_structs_ / batches_ : 0.0;
    out[10] = (double)peak_;
    out[11] = (double)budget_;
    out[12] = std::chrono::duration<double>(now - t0_).count();
    out[13] = max_running_;
    out[14] = refused_;
    out[15] = md_steps_;
    if (reset) {
        lat_eval_.clear();
        lat_md_.clear();
        atoms_eval_ = atom_steps_ = md_steps_ = batches_ = batch_structs_ = refused_ = 0;
        max_running_ = running_;
        peak_ = reserved_total();
        t0_ = now;
    }
}

}  // namespace hdnnp
