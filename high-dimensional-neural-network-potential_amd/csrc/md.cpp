// Host side of the on-device molecular dynamics (see md.hpp).
#include "md.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

namespace hdnnp {

#define MDCHK(x)                                                                               \
    do {                                                                                       \
        hipError_t e_ = (x);                                                                   \
        if (e_ != hipSuccess)                                                                  \
            throw Error(5, std::string("HIP error ") + hipGetErrorString(e_) + " at " #x);     \
    } while (0)

MdSystems::MdSystems(Engine &eng, int nsys, const int *nat, const int *Z, const double *pos, const double *mom,
                     const double *mass, const int *periodic, const double *lat, double dt, double skin)
    : eng_(eng), dt_(dt), skin_(skin)
{
    try {
        init(nsys, nat, Z, pos, mom, mass, periodic, lat);
    } catch (...) {
        release();   // a destructor does not run for a constructor that throws
        throw;
    }
}

void MdSystems::init(int nsys, const int *nat, const int *Z, const double *pos, const double *mom, const double *mass,
                     const int *periodic, const double *lat)
{
    if (nsys < 1 || !nat || !Z || !pos || !mom || !mass || !periodic) throw Error(1, "null argument or nsys < 1");
    if (!(dt_ > 0) || !(skin_ >= 0)) throw Error(1, "dt must be > 0 and skin >= 0");
    eng_.ensure_device();
    st_.resize(nsys);
    std::vector<int> off(nsys + 1, 0);
    long total = 0;
    for (int k = 0; k < nsys; ++k) {
        if (nat[k] < 0) throw Error(1, "nat < 0");
        st_[k].nat = nat[k];
        st_[k].periodic = periodic[k] ? 1 : 0;
        if (st_[k].periodic) {
            if (!lat) throw Error(1, "periodic system without lattice");
            std::memcpy(st_[k].lat, lat + 9 * (size_t)k, sizeof st_[k].lat);
        }
        total += nat[k];
        if (total >= (1L << 30)) throw Error(6, "too many atoms");
        off[k + 1] = (int)total;
    }
    nat_ = (int)total;
    const size_t n = std::max<size_t>(1, nat_);
    for (int i = 0; i < nat_; ++i)
        if (!(mass[i] > 0)) throw Error(1, "masses must be positive");
    Z_.ensure(4 * n);
    pos_.ensure(24 * n);
    mom_.ensure(24 * n);
    mass_.ensure(8 * n);
    F_.ensure(24 * n);
    Fold_.ensure(24 * n);
    posb_.ensure(24 * n);
    wposb_.ensure(sizeof(double4) * n);
    off_.ensure(4 * (nsys + 1));
    ekin_.ensure(8 * nsys);
    maxd2_.ensure(8);
    const size_t q = (size_t(256) << 20) / (72 * n);   // saved states that fit in 256 MiB
    maxk_ = (int)std::min<size_t>(kMaxChunk, q > 2 ? q - 1 : 1);
    snap_.ensure(72 * n * (size_t)(maxk_ + 1));
    d2_.ensure(8 * kMaxChunk);
    ctr_.ensure(256);
    if (const char *g = std::getenv("HDNNP_MD_GRAPHS")) graphs_ = std::strcmp(g, "0") != 0;
    {
        HipUnsafe guard;
        MDCHK(hipHostMalloc((void **)&d2_host_, 8 * kMaxChunk, hipHostMallocDefault));
    }
    const hipStream_t s = eng_.stream();          // copies ordered on it, not on the legacy stream
    MDCHK(hipMemcpyAsync(Z_.p, Z, 4 * (size_t)nat_, hipMemcpyHostToDevice, s));
    MDCHK(hipMemcpyAsync(pos_.p, pos, 24 * (size_t)nat_, hipMemcpyHostToDevice, s));
    MDCHK(hipMemcpyAsync(mom_.p, mom, 24 * (size_t)nat_, hipMemcpyHostToDevice, s));
    MDCHK(hipMemcpyAsync(mass_.p, mass, 8 * (size_t)nat_, hipMemcpyHostToDevice, s));
    MDCHK(hipMemcpyAsync(off_.p, off.data(), 4 * (size_t)(nsys + 1), hipMemcpyHostToDevice, s));
    MDCHK(hipStreamSynchronize(s));
    epot_.assign(nsys, 0.0);
    ekin_h_.assign(nsys, 0.0);
    forces(nullptr);     // initial forces (the reference's call before the loop)
}

MdSystems::~MdSystems() { release(); }

void MdSystems::release()
{
    HipUnsafe guard;
    for (DevMem *d : {&Z_, &pos_, &mom_, &mass_, &F_, &Fold_, &posb_, &wposb_, &off_, &ekin_, &maxd2_, &snap_, &d2_,
                      &ctr_})
        d->release();
    if (d2_host_) (void)hipHostFree(d2_host_);
    d2_host_ = nullptr;
    if (exec_) (void)hipGraphExecDestroy(exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    exec_ = nullptr;
    graph_ = nullptr;
}

size_t MdSystems::device_bytes() const
{
    size_t b = 0;
    for (const DevMem *d : {&Z_, &pos_, &mom_, &mass_, &F_, &Fold_, &posb_, &wposb_, &off_, &ekin_, &maxd2_,
                            &snap_, &d2_})
        b += d->bytes;
    return b;
}

// Forces of the current positions. With skin > 0 the last lists are reused
// while every atom is within skin/2 of where it was at their build.
void MdSystems::forces(double *epot)
{
    hipStream_t s = eng_.stream();
    bool rebuild = !built_ || skin_ == 0.0 || eng_.build_generation() != gen_ || !eng_.prepared();
    if (!rebuild) {
        MDCHK(hipMemsetAsync(maxd2_.p, 0, 8, s));
        launch_md_disp(pos_.as<double>(), posb_.as<double>(), nat_, maxd2_.as<unsigned long long>(), nullptr, s);
        unsigned long long bits = 0;
        MDCHK(hipMemcpyAsync(&bits, maxd2_.p, 8, hipMemcpyDeviceToHost, s));
        MDCHK(hipStreamSynchronize(s));
        double d2;
        std::memcpy(&d2, &bits, 8);
        rebuild = !(d2 <= 0.25 * skin_ * skin_);
    }
    for (int attempt = 0;; ++attempt) {
        if (rebuild) {
            eng_.device_bboxes(st_, pos_.as<double>(), s);
            eng_.build(st_, Z_.as<int>(), pos_.as<double>(), s, eng_.model().maxrc + skin_);
            MDCHK(hipMemcpyAsync(posb_.p, pos_.p, 24 * (size_t)nat_, hipMemcpyDeviceToDevice, s));
            MDCHK(hipMemcpyAsync(wposb_.p, eng_.wpos(), sizeof(double4) * (size_t)nat_, hipMemcpyDeviceToDevice, s));
            ++builds_;
            if (built_ && since_ > 0) interval_ = since_;
            since_ = 0;
            built_ = true;
            careful_ = false;
            gen_ = eng_.build_generation();
        } else {
            launch_md_refresh(eng_.wpos(), wposb_.as<double4>(), pos_.as<double>(), posb_.as<double>(), nat_, s);
        }
        MDCHK(hipGetLastError());
        if (eng_.finish(F_.as<double>(), s, epot_.data(), nullptr)) break;
        if (attempt >= 2) throw Error(6, "neighbour capacity did not converge");
        rebuild = true;   // capacity exceeded: build again with room for it
    }
    if (epot) std::copy(epot_.begin(), epot_.end(), epot);
}

void MdSystems::careful_step()
{
    hipStream_t s = eng_.stream();
    launch_md_drift(pos_.as<double>(), mom_.as<double>(), F_.as<double>(), Fold_.as<double>(), mass_.as<double>(), dt_,
                    nat_, s);
    forces(nullptr);
    launch_md_kick(mom_.as<double>(), F_.as<double>(), Fold_.as<double>(), dt_, nat_, s);
    MDCHK(hipGetLastError());
    ++steps_;
    ++since_;
}

void MdSystems::step_launch(bool timed)
{
    hipStream_t s = eng_.stream();
    const int *ctr = ctr_.as<int>();
    launch_md_begin(pos_.as<double>(), mom_.as<double>(), F_.as<double>(), Fold_.as<double>(), mass_.as<double>(), dt_,
                    posb_.as<double>(), eng_.wpos(), wposb_.as<double4>(), nat_, d2_.as<unsigned long long>(), ctr, s);
    eng_.finish_launch(F_.as<double>(), s, timed);
    launch_md_end(mom_.as<double>(), F_.as<double>(), Fold_.as<double>(), dt_, snap_.as<double>(), pos_.as<double>(), nat_,
                  ctr, s);
    launch_md_tick(ctr_.as<int>(), s);
}

bool MdSystems::replay(int n)
{
    if (!graphs_) return false;
    hipStream_t s = eng_.stream();
    std::vector<char> key = eng_.launch_key();
    if (!exec_ || key != key_) {
        // no other thread allocates, frees or copies synchronously through
        // this library while the capture runs (HipUnsafe)
        HipUnsafe::Capture excl;
        if (exec_) (void)hipGraphExecDestroy(exec_);
        if (graph_) (void)hipGraphDestroy(graph_);
        exec_ = nullptr;
        graph_ = nullptr;
        if (hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed) != hipSuccess) {
            graphs_ = false;
            return false;
        }
        bool ok = true;
        try {
            step_launch(false);
        } catch (const std::exception &) {
            ok = false;
        }
        hipGraph_t g = nullptr;
        ok = hipStreamEndCapture(s, &g) == hipSuccess && ok && g;
        if (ok) ok = hipGraphInstantiate(&exec_, g, nullptr, nullptr, 0) == hipSuccess;
        graph_ = g;
        if (!ok) {   // launch directly from now on
            if (exec_) (void)hipGraphExecDestroy(exec_);
            if (graph_) (void)hipGraphDestroy(graph_);
            exec_ = nullptr;
            graph_ = nullptr;
            graphs_ = false;
            (void)hipGetLastError();
            return false;
        }
        key_ = std::move(key);
        ++captures_;
    }
    for (int i = 0; i < n; ++i) MDCHK(hipGraphLaunch(exec_, s));
    replays_ += n;
    return true;
}

int MdSystems::chunk(int k)
{
    hipStream_t s = eng_.stream();
    const size_t slot = 9 * (size_t)nat_;    // doubles per saved state
    double *sp = snap_.as<double>();
    launch_md_save(sp, pos_.as<double>(), mom_.as<double>(), F_.as<double>(), nat_, nullptr, s);
    MDCHK(hipMemsetAsync(d2_.p, 0, 8 * (size_t)k, s));
    MDCHK(hipMemsetAsync(ctr_.p, 0, 4, s));
    step_launch(true);                                   // sizes every buffer the step uses
    if (k > 1 && !replay(k - 1))
        for (int i = 1; i < k; ++i) step_launch(true);
    MDCHK(hipGetLastError());
    MDCHK(hipMemcpyAsync(d2_host_, d2_.p, 8 * (size_t)k, hipMemcpyDeviceToHost, s));
    const bool ok = eng_.finish_collect(s, epot_.data(), nullptr);   // waits for the stream
    const double lim = 0.25 * skin_ * skin_;
    int j = ok ? k : 0;              // first step whose lists were stale
    for (int i = 0; i < k && ok; ++i) {
        double d;
        std::memcpy(&d, &d2_host_[i], 8);
        if (!(d <= lim)) {
            j = i;
            break;
        }
    }
    if (j < k) {   // back to the state after step j-1 (the chunk's start for j = 0)
        launch_md_load(sp + j * slot, pos_.as<double>(), mom_.as<double>(), F_.as<double>(), nat_, s);
        MDCHK(hipGetLastError());
        ++rollbacks_;
    }
    steps_ += j;
    since_ += j;
    return j;
}

void MdSystems::step(int nsteps, double *epot, double *ekin)
{
    if (nsteps < 0) throw Error(1, "nsteps < 0");
    hipStream_t s = eng_.stream();
    for (int t = 0; t < nsteps;) {
        const bool reuse = skin_ > 0 && built_ && !careful_ && eng_.build_generation() == gen_ && eng_.prepared();
        if (!reuse) {
            careful_step();
            ++t;
            continue;
        }
        // through the step after the one the last interval predicts for the
        // next build, so that the build step usually falls inside the chunk
        const int want = interval_ > 0 ? interval_ - since_ + 1 : 4;
        const int k = std::min(nsteps - t, std::max(2, std::min(want, maxk_)));
        const int done = chunk(k);
        t += done;
        if (done < k) careful_ = true;   // one step at a time up to the list build
    }
    launch_md_ekin(mom_.as<double>(), mass_.as<double>(), off_.as<int>(), (int)st_.size(), ekin_.as<double>(), s);
    MDCHK(hipMemcpyAsync(ekin_h_.data(), ekin_.p, 8 * st_.size(), hipMemcpyDeviceToHost, s));
    MDCHK(hipStreamSynchronize(s));
    if (epot) std::copy(epot_.begin(), epot_.end(), epot);
    if (ekin) std::copy(ekin_h_.begin(), ekin_h_.end(), ekin);
}

void MdSystems::get(double *pos, double *mom, double *force)
{
    hipStream_t s = eng_.stream();
    if (pos) MDCHK(hipMemcpyAsync(pos, pos_.p, 24 * (size_t)nat_, hipMemcpyDeviceToHost, s));
    if (mom) MDCHK(hipMemcpyAsync(mom, mom_.p, 24 * (size_t)nat_, hipMemcpyDeviceToHost, s));
    if (force) MDCHK(hipMemcpyAsync(force, F_.p, 24 * (size_t)nat_, hipMemcpyDeviceToHost, s));
    MDCHK(hipStreamSynchronize(s));
}

}  // namespace hdnnp
