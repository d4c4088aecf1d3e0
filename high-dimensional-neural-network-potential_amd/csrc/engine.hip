// Host orchestration of one HDNNP evaluation on the GPU.
//
// Reference flow: hdnnpEnergyAndForces (src/potentials/hdnnps.f90:333-371)
// = neighbour lists (:373-388) -> symmetry functions (:390-485) -> NNs,
// forces and dEdLat (:487-562). Here:
//   map elements -> wrap + bin -> cell sort -> neighbour count / scan / fill
//   -> partner entries -> SF values -> MLP (+dE/dG) -> SF gradients
//   -> force gather -> FP64 energy / virial reductions.
// This file is built with -ffp-contract=off: the cell-grid parameters below
// must be rounded exactly as in the reference.
#include "engine.hpp"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

namespace hdnnp {

#pragma clang fp contract(off)

#define HIPCHK(x)                                                                              \
    do {                                                                                       \
        hipError_t e_ = (x);                                                                   \
        if (e_ != hipSuccess)                                                                  \
            throw Error(5, std::string("HIP error ") + hipGetErrorString(e_) + " at " #x);     \
    } while (0)

std::shared_mutex &capture_mutex()
{
    static std::shared_mutex mu;
    return mu;
}

static thread_local int t_unsafe_depth = 0;     // nested HipUnsafe scopes of this thread
static thread_local bool t_capturing = false;   // this thread holds the mutex for a capture

HipUnsafe::HipUnsafe()
{
    if (t_capturing) return;
    counted_ = true;
    if (t_unsafe_depth++ == 0) {
        capture_mutex().lock_shared();
        locked_ = true;
    }
}

HipUnsafe::~HipUnsafe()
{
    if (locked_) capture_mutex().unlock_shared();
    if (counted_) --t_unsafe_depth;
}

HipUnsafe::Capture::Capture() : lk(capture_mutex()) { t_capturing = true; }
HipUnsafe::Capture::~Capture() { t_capturing = false; }

// Work buffers of Engine::evaluate_batch_device_noalloc: carved from this
// array instead of allocated. Every device the library runs on holds its own
// copy until the process ends. In host code the name stands for no device
// address: host code takes the address per device from hipGetSymbolAddress
// (scratch_address) and uses only that.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// the scratch address on device dev (resolved once per device, with it current)
static void *scratch_address(int dev)
{
    constexpr int kMaxDevices = 64;
    static std::mutex mu;
    static void *addr[kMaxDevices] = {};
    if (dev < 0 || dev >= kMaxDevices) throw Error(5, "device index out of range");
    std::lock_guard<std::mutex> lk(mu);
    if (!addr[dev]) {
        HipUnsafe guard;
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw Error(5, std::string("hipGetSymbolAddress of the device scratch failed: ") + hipGetErrorString(e));
        addr[dev] = p;
    }
    return addr[dev];
}

// carving state of a no-allocation call on this thread (null otherwise)
struct Arena {
    char *base;
    size_t size, used;
};
static thread_local Arena *t_arena = nullptr;

void DevMem::ensure(size_t b)
{
    if (b <= bytes && p) return;
    if (t_arena) {
        const size_t need = (std::max<size_t>(b, 1) + 255) / 256 * 256;
        if (need > t_arena->size - t_arena->used)
            throw Error(6, "the evaluation needs more than the " + std::to_string(t_arena->size >> 20) +
                               " MiB of device scratch of the entry point that allocates nothing");
        p = t_arena->base + t_arena->used;
        t_arena->used += need;
        bytes = need;
        return;
    }
    HipUnsafe guard;
    release();
    const size_t want = std::max<size_t>(256, b + b / 4);
    HIPCHK(hipMalloc(&p, want));
    bytes = want;
}

void DevMem::release()
{
    if (t_arena) {   // carved from the scratch: nothing to free
        p = nullptr;
        bytes = 0;
        return;
    }
    if (p) {
        HipUnsafe guard;
        (void)hipFree(p);
    }
    p = nullptr;
    bytes = 0;
}

// det3D and inv3DM (src/math/linalg.f90:227-238, 264-289)
double det3(const double *M)
{
#define A3(r, c) M[(r) + 3 * (c)]
    double t = A3(0, 0) * A3(1, 1) * A3(2, 2);
    t = t + A3(0, 1) * A3(1, 2) * A3(2, 0);
    t = t + A3(0, 2) * A3(1, 0) * A3(2, 1);
    t = t - A3(0, 0) * A3(1, 2) * A3(2, 1);
    t = t - A3(0, 1) * A3(1, 0) * A3(2, 2);
    t = t - A3(0, 2) * A3(1, 1) * A3(2, 0);
    return t;
#undef A3
}

void inv3(const double *A, double *B)
{
#define A3(r, c) A[(r) + 3 * (c)]
    const double a1[3] = {A3(0, 0), A3(0, 1), A3(0, 2)};
    const double a2[3] = {A3(1, 0), A3(1, 1), A3(1, 2)};
    const double a3[3] = {A3(2, 0), A3(2, 1), A3(2, 2)};
#undef A3
    const double det = det3(A);
    const double b1[3] = {a2[1] * a3[2] - a2[2] * a3[1], a1[2] * a3[1] - a1[1] * a3[2],
                          a1[1] * a2[2] - a1[2] * a2[1]};
    const double b2[3] = {a2[2] * a3[0] - a2[0] * a3[2], a1[0] * a3[2] - a1[2] * a3[0],
                          a1[2] * a2[0] - a1[0] * a2[2]};
    const double b3[3] = {a2[0] * a3[1] - a2[1] * a3[0], a1[1] * a3[0] - a1[0] * a3[1],
                          a1[0] * a2[1] - a1[1] * a2[0]};
    for (int c = 0; c < 3; ++c) {
        B[0 + 3 * c] = b1[c] / det;
        B[1 + 3 * c] = b2[c] / det;
        B[2 + 3 * c] = b3[c] / det;
    }
}

// getNcells (src/atoms/pointsets.f90:220-243)
static void get_ncells(const double *L, double rc, int n[3])
{
    auto cross = [](const double *a, const double *b, double *c) {
        c[0] = a[1] * b[2] - a[2] * b[1];
        c[1] = a[2] * b[0] - a[0] * b[2];
        c[2] = a[0] * b[1] - a[1] * b[0];
    };
    double axb[3], axc[3], bxc[3];
    cross(L + 0, L + 3, axb);
    cross(L + 0, L + 6, axc);
    cross(L + 3, L + 6, bxc);
    for (double *v : {axb, axc, bxc}) {
        const double nr = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        for (int k = 0; k < 3; ++k) v[k] = v[k] / nr;
    }
    const double p0 = L[0] * bxc[0] + L[1] * bxc[1] + L[2] * bxc[2];
    const double p1 = L[3] * axc[0] + L[4] * axc[1] + L[5] * axc[2];
    const double p2 = L[6] * axb[0] + L[7] * axb[1] + L[8] * axb[2];
    n[0] = (int)std::ceil(rc / std::fabs(p0));
    n[1] = (int)std::ceil(rc / std::fabs(p1));
    n[2] = (int)std::ceil(rc / std::fabs(p2));
}

// ---------------------------------------------------------------------------

Engine::Engine(std::vector<HostModel> members, int precision) : precision_(precision)
{
    if (members.empty()) throw Error(1, "no potential given");
    model_ = std::move(members[0]);
    members_.assign(std::make_move_iterator(members.begin() + 1), std::make_move_iterator(members.end()));
    if (precision_ < 0 || precision_ > 2)
        throw Error(1, "precision must be HDNNP_FP64, HDNNP_MIXED or HDNNP_MIXED_BF16");
    if ((int)model_.el.size() > kMaxEl) throw Error(6, "too many elements for this build");
    if (const char *v = std::getenv("HDNNP_EVAL_GRAPHS")) eval_graphs_on_ = std::strcmp(v, "1") == 0;
    if (const char *v = std::getenv("HDNNP_STAGE_TIMINGS")) stage_timing_ = std::strcmp(v, "1") == 0;
    // committee members (globalHdnnps.f90:31-70) reuse the first potential's
    // symmetry functions and scaling and differ in their networks and atom
    // energies: every element of the first needs a network with as many inputs
    for (size_t k = 0; k < members_.size(); ++k) {
        const HostModel &mk = members_[k];
        for (const ElementModel &em : model_.el) {
            const int r = mk.rank_of_Z(em.Z);
            if (r < 0) throw Error(3, "committee member " + std::to_string(k + 2) + " has no network for Z=" +
                                          std::to_string(em.Z));
            const ElementModel &ek = mk.el[r];
            if (ek.layers.empty() || ek.layers[0].nin != (int)em.sfs.size())
                throw Error(3, "committee member " + std::to_string(k + 2) + ": network of Z=" +
                                   std::to_string(em.Z) + " does not take the first potential's " +
                                   std::to_string(em.sfs.size()) + " symmetry functions");
            if (ek.layers.size() != mk.el[0].layers.size())
                throw Error(3, "committee member " + std::to_string(k + 2) + ": networks differ in depth");
        }
    }
    for (size_t e = 0; e < model_.el.size(); ++e) {
        elem_valid_[e] = model_.el[e].valid;
        elem_why_[e] = model_.el[e].why;
    }
}

std::vector<DevMem *> Engine::work_buffers()
{
    return {&bbox_, &geo_, &sid_, &csid_, &chunks_, &sc_, &res_, &partial_,
            &posin_, &fout_, &eidx_, &wpos_, &cellkey_, &cellkey_sorted_, &cell_atoms_,
            &iota_, &cell_start_, &cell_end_, &count_, &perm_, &Ei_,
            &vir_, &ent_j_, &ent_shift_, &ent_s_, &rev_, &pos_start_, &gent_, &G_, &dEdG_, &cub_tmp_, &elem_chunks_,
            &trip_, &tcount_, &sfpar_, &sfpar_off_, &stage_G_, &stage_dsf_, &stage_off_};
}

size_t Engine::device_bytes() const
{
    size_t b = 0;
    for (DevMem *d : const_cast<Engine *>(this)->work_buffers()) b += d->bytes;
    return b;
}

void Engine::release_work()
{
    if (stream_) HIPCHK(hipStreamSynchronize(stream_));
    drop_eval_graphs();
    for (DevMem *d : work_buffers()) d->release();
    uploaded_.clear();
    prep_ = Prepared{};
    ++build_gen_;
    last_nat_ = 0;
    last_total_ = 0;
}

Engine::~Engine()
{
    HipUnsafe guard;
    if (stream_) (void)hipStreamSynchronize(stream_);
    drop_eval_graphs();
    for (auto &d : model_mem_) d.release();
    for (DevMem *d : work_buffers()) d->release();
    misc_.release();
    if (pinned_) (void)hipHostFree(pinned_);
    if (io_pinned_) (void)hipHostFree(io_pinned_);
    for (auto &e : ev_)
        if (e) (void)hipEventDestroy(e);
    if (stream_) (void)hipStreamDestroy(stream_);
}

template <typename T>
static const T *upload(std::vector<DevMem> &keep, const std::vector<T> &v)
{
    keep.emplace_back();
    keep.back().ensure(sizeof(T) * std::max<size_t>(1, v.size()));
    HipUnsafe guard;
    if (!v.empty()) HIPCHK(hipMemcpy(keep.back().p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
    return keep.back().as<T>();
}

void Engine::init_device()
{
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) throw Error(5, "no HIP device available");
    {
        HipUnsafe guard;
        HIPCHK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
        for (auto &e : ev_) HIPCHK(hipEventCreate(&e));
    }
    HIPCHK(hipGetDevice(&dev_));
    scratch_ = scratch_address(dev_);

    const HostModel &m = model_;
    DevModel d{};
    d.nel = (int)m.el.size();
    d.fallback_el = -1;
    d.cutoff_type = m.cutoff_type;
    std::vector<int> zlut(128, -1);
    std::vector<double> mean, den;
    std::vector<int2> g2tab(kMaxEl * kMaxEl, make_int2(0, 0)), g4tab(kMaxEl * kMaxEl * kMaxEl, make_int2(0, 0));
    std::vector<int> g2s, g4s, g4ipow, g4sq, g4new;
    std::vector<double> g2eta, g2rs, g4eta, g4lam, g4zeta, g4pre;
    int off = 0;
    for (int e = 0; e < d.nel; ++e) {
        const ElementModel &em = m.el[e];
        if (em.Z < 128) zlut[em.Z] = em.valid ? e : -2 - e;
        if (em.valid && d.fallback_el < 0) d.fallback_el = e;
        d.nsf[e] = (int)em.sfs.size();
        d.sf_off[e] = off;
        off += d.nsf[e];
        d.nsfmax = std::max(d.nsfmax, d.nsf[e]);
        d.rc[e] = em.sfs.empty() ? 0.0 : em.sfs[0].rcut;
        d.atom_energy[e] = em.atom_energy;
        for (auto &s : em.sfs) {
            // scaled G = (G - c) / den: the reference's centring and scaling
            // (hdnnps.f90:471-481), or RuNNer's min-max form when
            // scale_min_short / scale_max_short are given
            double c = m.centersf ? s.sfmean : 0.0, dn = m.scalesf ? s.sfmax - s.sfmin : 1.0;
            if (m.scale_range && m.scalesf) {
                dn = (s.sfmax - s.sfmin) / (m.scale_max - m.scale_min);
                c = (m.centersf ? s.sfmean : s.sfmin) - m.scale_min * dn;
            }
            mean.push_back(c);
            den.push_back(dn);
        }
        for (int f = 0; f < d.nel; ++f) {          // G2 of (e, f)
            const int start = (int)g2s.size();
            for (int s = 0; s < d.nsf[e]; ++s) {
                const SymFn &sf = em.sfs[s];
                if (sf.type == 2 && sf.e1 == m.el[f].Z) {
                    g2s.push_back(s);
                    g2eta.push_back(sf.eta);
                    g2rs.push_back(sf.rs);
                }
            }
            g2tab[e * kMaxEl + f] = make_int2(start, (int)g2s.size() - start);
        }
        for (int f = 0; f < d.nel; ++f)               // G4 of (e, {f, g})
            for (int g = f; g < d.nel; ++g) {
                const int start = (int)g4s.size();
                double last = 0.0;
                for (int s = 0; s < d.nsf[e]; ++s) {
                    const SymFn &sf = em.sfs[s];
                    if (sf.type != 3) continue;
                    const int za = m.el[f].Z, zb = m.el[g].Z;
                    if (!((sf.e1 == za && sf.e2 == zb) || (sf.e1 == zb && sf.e2 == za))) continue;
                    const int iz = (int)sf.zeta;
                    const bool integral = std::fabs(sf.zeta - iz) < 1.e-10;
                    double pre;
                    if (integral) {    // 2**(1-intzeta) by binary powering, exact
                        pre = std::ldexp(1.0, 1 - iz);
                    } else {
                        pre = std::pow(2.0, 1.0 - sf.zeta);
                    }
                    const bool first = (int)g4s.size() == start;
                    g4new.push_back(first || std::fabs(sf.eta - last) > 1.e-10 ? 1 : 0);
                    if (g4new.back()) last = sf.eta;
                    g4s.push_back(s);
                    g4eta.push_back(g4new.back() ? sf.eta : last);
                    g4lam.push_back(sf.lambda);
                    g4zeta.push_back(sf.zeta);
                    g4pre.push_back(pre);
                    g4ipow.push_back(integral ? iz - 1 : kNonIntegralZeta);
                    int sq = -1;
                    for (int k = 0; k < 31 && integral; ++k)
                        if (iz == (1 << k)) sq = k;
                    g4sq.push_back(sq);
                }
                g4tab[(e * kMaxEl + f) * kMaxEl + g] = make_int2(start, (int)g4s.size() - start);
            }
    }
    {   // per-element ranges, neighbour element of each G2 entry, power bits
        std::vector<int> g2f(g2s.size(), 0);
        for (int e = 0; e < d.nel; ++e) {
            d.g2_start[e] = g2tab[e * kMaxEl].x;
            d.g4_start[e] = g4tab[(e * kMaxEl) * kMaxEl].x;
            int n2 = 0, n4 = 0;
            for (int f = 0; f < d.nel; ++f) {
                const int2 t = g2tab[e * kMaxEl + f];
                for (int k = 0; k < t.y; ++k) g2f[t.x + k] = f;
                n2 += t.y;
                for (int g = f; g < d.nel; ++g) n4 += g4tab[(e * kMaxEl + f) * kMaxEl + g].y;
            }
            d.g2_count[e] = n2;
            d.g4_count[e] = n4;
            d.g4_max = std::max(d.g4_max, n4);
        }
        d.g2_f = upload(model_mem_, g2f);
        int maxp = 0;
        for (int ip : g4ipow)
            if (ip != kNonIntegralZeta) maxp = std::max(maxp, std::abs(ip));
        d.powbits = 0;
        while ((1 << d.powbits) <= maxp) ++d.powbits;
        d.g4_lampm = 1;
        for (size_t k = 0; k < g4lam.size(); ++k)
            if (!(g4lam[k] == 1.0 || g4lam[k] == -1.0) || !(g4zeta[k] >= 1.0)) d.g4_lampm = 0;
        d.g4_nblk = 0;
        for (int e = 0; e < d.nel; ++e) {
            int nb = 0;
            for (int f = 0; f < d.nel; ++f)
                for (int g = f; g < d.nel; ++g) nb += g4tab[(e * kMaxEl + f) * kMaxEl + g].y > 0;
            d.g4_nblk = std::max(d.g4_nblk, nb);
        }
        auto tof = [](const std::vector<double> &v) { return std::vector<float>(v.begin(), v.end()); };
        d.g2f_eta = upload(model_mem_, tof(g2eta));
        d.g2f_rs = upload(model_mem_, tof(g2rs));
        d.g4f_eta = upload(model_mem_, tof(g4eta));
        d.g4f_lam = upload(model_mem_, tof(g4lam));
        d.g4f_zeta = upload(model_mem_, tof(g4zeta));
        d.g4f_pre = upload(model_mem_, tof(g4pre));
        auto toea = [](const std::vector<double> &v) {
            std::vector<float> o(v.size());
            for (size_t k = 0; k < v.size(); ++k) o[k] = (float)(-v[k] * 1.4426950408889634);
            return o;
        };
        d.g2f_ea = upload(model_mem_, toea(g2eta));
        d.g4f_ea = upload(model_mem_, toea(g4eta));
        std::vector<ElemInfo> ei(std::max(d.nel, 1));
        for (int e = 0; e < d.nel; ++e) {
            ei[e].nsf = d.nsf[e];
            ei[e].sf_off = d.sf_off[e];
            ei[e].g2_start = d.g2_start[e];
            ei[e].g2_count = d.g2_count[e];
            ei[e].g4_start = d.g4_start[e];
            ei[e].g4_count = d.g4_count[e];
            ei[e].rc = d.rc[e];
            ei[e].atom_energy = d.atom_energy[e];
        }
        d.einfo = upload(model_mem_, ei);
        ei_host_ = ei;
    }
    d.zlut = upload(model_mem_, zlut);
    d.sf_mean = upload(model_mem_, mean);
    d.sf_den = upload(model_mem_, den);
    d.g2_tab = upload(model_mem_, g2tab);
    d.g4_tab = upload(model_mem_, g4tab);
    d.g2_s = upload(model_mem_, g2s);
    d.g2_eta = upload(model_mem_, g2eta);
    d.g2_rs = upload(model_mem_, g2rs);
    d.g4_s = upload(model_mem_, g4s);
    d.g4_eta = upload(model_mem_, g4eta);
    d.g4_lam = upload(model_mem_, g4lam);
    d.g4_zeta = upload(model_mem_, g4zeta);
    d.g4_pre = upload(model_mem_, g4pre);
    d.g4_ipow = upload(model_mem_, g4ipow);
    d.g4_sq = upload(model_mem_, g4sq);
    d.g4_neweta = upload(model_mem_, g4new);
    {   // G5 (type 9) functions per centre element
        std::vector<SfParam> g5;
        for (int e = 0; e < d.nel; ++e) {
            d.g5_start[e] = (int)g5.size();
            for (int s = 0; s < d.nsf[e]; ++s) {
                const SymFn &sf = m.el[e].sfs[s];
                if (sf.type != 9) continue;
                SfParam p{};
                p.type = 9;
                p.f = m.rank_of_Z(sf.e1);
                p.g = m.rank_of_Z(sf.e2);
                p.s = s;
                p.eta = sf.eta;
                p.lambda = sf.lambda;
                p.zeta = sf.zeta;
                p.pre = std::pow(2.0, 1.0 - sf.zeta);
                g5.push_back(p);
            }
            d.g5_count[e] = (int)g5.size() - d.g5_start[e];
        }
        d.g5_total = (int)g5.size();
        d.g5 = upload(model_mem_, g5);
    }
    d.nlayers = m.el.empty() ? 0 : (int)m.el[0].layers.size();
    mlp_maxw_ = d.nsfmax;
    mlp_sumc_ = 0;
    for (int l = 0; l < d.nlayers; ++l) {
        d.nodes[l] = m.el[0].layers[l].nout;
        d.act[l] = m.el[0].layers[l].act;
        mlp_maxw_ = std::max(mlp_maxw_, d.nodes[l]);
        mlp_sumc_ += d.nodes[l];
    }
    for (int e = 0; e < d.nel; ++e) {
        std::vector<double> w;
        for (auto &L : m.el[e].layers) {          // W(i,j) at i + nout*j == [j][i]
            w.insert(w.end(), L.W.begin(), L.W.end());
            w.insert(w.end(), L.b.begin(), L.b.end());
        }
        std::vector<float> wf(w.begin(), w.end());
        d.W64[e] = upload(model_mem_, w);
        d.W32[e] = upload(model_mem_, wf);
    }
    dm_ = d;
    dmm_.assign(1, d);
    maxw_m_.assign(1, mlp_maxw_);
    // committee members: the first potential's tables with the member's networks
    for (const HostModel &mk : members_) {
        DevModel dk = d;
        std::vector<ElemInfo> eik = ei_host_;
        dk.nlayers = (int)mk.el[0].layers.size();
        if (dk.nlayers > kMaxLay) throw Error(6, "network too deep for this build");
        int maxw = d.nsfmax;
        for (int l = 0; l < dk.nlayers; ++l) {
            dk.nodes[l] = mk.el[0].layers[l].nout;
            dk.act[l] = mk.el[0].layers[l].act;
            maxw = std::max(maxw, dk.nodes[l]);
        }
        for (int e = 0; e < d.nel; ++e) {
            const ElementModel &ek = mk.el[mk.rank_of_Z(m.el[e].Z)];
            eik[e].atom_energy = ek.atom_energy;
            dk.atom_energy[e] = ek.atom_energy;
            std::vector<double> w;
            for (auto &L : ek.layers) {
                w.insert(w.end(), L.W.begin(), L.W.end());
                w.insert(w.end(), L.b.begin(), L.b.end());
            }
            std::vector<float> wf(w.begin(), w.end());
            dk.W64[e] = upload(model_mem_, w);
            dk.W32[e] = upload(model_mem_, wf);
        }
        dk.einfo = upload(model_mem_, eik);
        dmm_.push_back(dk);
        maxw_m_.push_back(maxw);
    }
    misc_.ensure(1024);
    dev_ready_ = true;
}

// misc buffer layout
static constexpr int kIntElemCount = 0, kIntElemOff = 16, kIntFlags = 40;   // ints
static constexpr int kIntReduceDone = 56;   // the reduction's count of finished blocks
static constexpr size_t kFlagBytes = 8 * sizeof(int);                          // w.flags[0..8)
// host-array evaluations of at most this many atoms stage through pinned memory
static constexpr long kSmallHostEval = 4096;
static constexpr size_t kDblOffset = 256;                                    // bytes
// largest per-(atom, stencil cell) slot table; beyond it partners are found by bisection
static constexpr size_t kPosStartBytes = size_t(1) << 30;
// atoms per block of the first reduction stage
static constexpr int kReduceChunk = 512;

template <typename R>
void Engine::compute(Work &w, const Geo &g, hipStream_t s, bool timed)
{
    const int nat = w.nat;
    G_.ensure(sizeof(R) * (size_t)nat * dm_.nsfmax);
    dEdG_.ensure(sizeof(R) * (size_t)nat * dm_.nsfmax);
    gent_.ensure(sizeof(R) * 3 * (size_t)nat * w.cap);
    R *G = G_.as<R>(), *dEdG = dEdG_.as<R>(), *gent = gent_.as<R>();

    const int ga = sf_tile(sizeof(R) == 8, w.lim, tile_, dm_.nsfmax, dm_.g4_max, dm_.nel);
    // one atom per wavefront: pass A hands its compacted triplets to pass B
    if (ga == 1) {
        trip_.ensure(4 * (size_t)nat * w.tcap);
        tcount_.ensure(4 * (size_t)nat * (1 + kTripBlk));
        w.trip = trip_.as<int>();
        w.tcount = tcount_.as<int>();
        w.tblk = w.tcount + nat;
    } else {
        w.trip = nullptr;
        w.tcount = nullptr;
        w.tblk = nullptr;
    }
    if (!skip_values_) {
        launch_sf_values<R>(dm_, g, w, w.lim, ga, G, s);   // clears w.tblk unless it writes the tables
        if (dm_.g5_total) launch_g5_values<R>(dm_, g, w, G, s);
        HIPCHK(hipGetLastError());
        trip_valid_ = w.trip != nullptr;
        tblk_valid_ = w.tblk != nullptr;
    } else if (!trip_valid_) {   // G came from elsewhere: pass B compacts its own triplets
        w.trip = nullptr;
        w.tcount = nullptr;
        w.tblk = nullptr;
    } else if (!tblk_valid_) {
        w.tblk = nullptr;
    }
    if (timed) HIPCHK(hipEventRecord(ev_[3], s));

    // one block per 16-atom tile of one element; the element counts stay on
    // the device, so the grid covers the most tiles any split can need and
    // surplus blocks exit at once. A committee runs every member's network on
    // the same G: each writes its atomic energies, dE/dG is averaged (forces
    // and dE/dh are linear in it).
    const int nm = (int)dmm_.size();
    for (int k = 0; k < nm; ++k) {
        if (mlp_lds_bytes(sizeof(R) == 8, dmm_[k], maxw_m_[k]) > 65536) throw Error(6, "MLP too wide for this build");
        launch_mlp_t<R>(dmm_[k], w, maxw_m_[k], (nat + 15) / 16 + dm_.nel, G, dEdG, w.Ei + (size_t)k * nat,
                        R(1) / R(nm), k > 0, s, precision_ == 2);
    }
    HIPCHK(hipGetLastError());
    if (timed) HIPCHK(hipEventRecord(ev_[4], s));

    launch_sf_grad<R>(dm_, g, w, w.lim, ga, dEdG, gent, s);
    if (dm_.g5_total) launch_g5_grad<R>(dm_, g, w, dEdG, gent, s);
    HIPCHK(hipGetLastError());
    if (timed) HIPCHK(hipEventRecord(ev_[5], s));

    launch_force_gather_t<R>(w, gent, s);
    HIPCHK(hipGetLastError());
}

// pinned host staging, grown on demand (only touched between synchronised runs)
static void *grow_pinned(void *&p, size_t &cap, size_t bytes)
{
    if (bytes > cap) {
        HipUnsafe guard;
        if (p) HIPCHK(hipHostFree(p));
        p = nullptr;
        cap = 0;
        const size_t want = std::max<size_t>(4096, bytes + bytes / 2);
        HIPCHK(hipHostMalloc(&p, want, hipHostMallocDefault));
        cap = want;
    }
    return p;
}
void *Engine::staging(size_t bytes)
{
    if (t_arena) {   // a no-allocation call stages through pageable memory
        if (host_stage_.size() < bytes) host_stage_.resize(bytes);
        stage_ptr_ = host_stage_.data();
    } else {
        stage_ptr_ = grow_pinned(pinned_, pinned_bytes_, bytes);
    }
    return stage_ptr_;
}
void *Engine::io_staging(size_t bytes) { return grow_pinned(io_pinned_, io_pinned_bytes_, bytes); }

// Cell grid of one structure (neighborLists.f90:131-169): lattice (periodic)
// or artificial box from the bounding box (nonperiodic), grid cells of at
// least 0.6 rc, stencil half-widths from getNcells of the grid.
static void structure_grid(const StructIn &in, double rc, StructGeo &G)
{
    std::memset(&G, 0, sizeof G);
    G.periodic = in.periodic ? 1 : 0;
    if (in.periodic) {
        std::memcpy(G.lat, in.lat, sizeof G.lat);
    } else {                                     // neighborLists.f90:131-138
        for (int a = 0; a < 3; ++a) {
            G.origin[a] = in.bbox[a] - 1.e-4;
            G.lat[a + 3 * a] = (in.bbox[3 + a] - G.origin[a]) + 1.e-4;
        }
    }
    inv3(G.lat, G.inv);
    double grid[9];
    const double target = rc * 0.6;               // neighborLists.f90:142-146
    for (int a = 0; a < 3; ++a) {
        const double *v = G.lat + 3 * a;
        const double len = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        if (!(len > 0) || !std::isfinite(len)) throw Error(1, "degenerate lattice");
        const double ngd = std::floor(len / target) + 1;
        if (ngd > 1e6) throw Error(6, "cell grid too large");
        G.ng[a] = (int)ngd;
        for (int k = 0; k < 3; ++k) grid[3 * a + k] = v[k] / G.ng[a];
    }
    get_ncells(grid, rc, G.nc);                   // neighborLists.f90:169
    // nonperiodic: offsets outside the grid are skipped anyway, so the stencil
    // is clamped to the grid (same cells visited; flat inputs stay cheap)
    if (!in.periodic)
        for (int a = 0; a < 3; ++a) G.nc[a] = std::min(G.nc[a], G.ng[a] - 1);
    const long ncell = (long)G.ng[0] * G.ng[1] * G.ng[2];
    if (ncell > (1L << 26)) throw Error(6, "cell grid too large");
    const long S = (long)(2 * G.nc[0] + 1) * (2 * G.nc[1] + 1) * (2 * G.nc[2] + 1);
    if (S > (1L << 22)) throw Error(6, "neighbour stencil too large");
    G.ncell = (int)ncell;
    G.S = (int)S;
    G.nat = in.nat;
}

// neighbour estimate of one structure from its mean density (as build uses it)
static double neighbour_estimate(const StructGeo &G, double rc)
{
    if (G.nat == 0) return 0.0;
    const double vol = std::fabs(det3(G.lat));
    double e = 4.0 / 3.0 * M_PI * rc * rc * rc * G.nat / std::max(vol, 1e-300);
    if (!G.periodic) e = std::min(e, (double)(G.nat - 1));
    return e;
}

// entry capacity per atom: a power of two from 64 covering the estimate (or
// the most neighbours this engine has seen), at most lim
// staging sizes of the symmetry-function kernels (entries per atom)
static const int kStage[] = {64, 96, 128, 256, 512, 1024};

// the smallest staging size that holds `want` entries, at most `lim`
static int entry_capacity(double est, int hint, int lim)
{
    const double want = std::max((double)hint, std::min(1.2 * est + 8.0, 1e6));
    int cap = kStage[0];
    for (int c : kStage) {
        cap = c;
        if (cap >= want || cap >= lim) break;
    }
    return cap;
}

// Triplets per atom the value pass keeps for the gradient pass: the pairs of
// lim neighbours, at most kTripletSlots. An atom with more falls back to the
// gradient pass's own compaction: the same triplets in the same order, but
// in other chunks, so that its sums round differently.
// Whether an atom falls back therefore must not depend on the evaluation it
// is part of (a batch result equals the structure's result alone, bit for
// bit): an atom has at most n(n-1)/2 <= lim(lim-1)/2 triplets, so the test
// count <= min(pairs(lim), kTripletSlots) is count <= kTripletSlots for every
// lim. 2048 holds the ~700-1,330 triplets of every perovskite atom.
constexpr int kTripletSlots = 2048;
static int triplet_capacity(int lim)
{
    const long pairs = (long)lim * (lim - 1) / 2;
    return (int)std::min<long>(pairs, kTripletSlots);
}

// the slot stride of a list holding `lim` entries per atom: a power of two
static int slot_stride(int lim)
{
    int c = 64;
    while (c < lim) c *= 2;
    return c;
}

void Engine::footprint_add(Footprint &f, const StructIn &in, double rc_list) const
{
    StructGeo G;
    structure_grid(in, rc_list, G);
    f.nat += in.nat;
    f.ncell += G.ncell;
    f.ps += (long)in.nat * G.S;
    f.nstruct += 1;
    f.est = std::max(f.est, neighbour_estimate(G, rc_list));
}

size_t Engine::footprint_bytes(const Footprint &f, bool slack) const
{
    const size_t R = precision_ == 0 ? 8 : 4, nm = (size_t)committee_size();
    size_t nsfmax = 0;
    for (const ElementModel &em : model_.el) nsfmax = std::max(nsfmax, em.sfs.size());
    const size_t cap = (size_t)slot_stride(entry_capacity(f.est, cap_hint_, 1024));
    const size_t nat = (size_t)f.nat;
    size_t b = nat * (9 * 4 + 32 + 8 * nm + 72 + 4 + 4 + 24 + 24 + 2 * R * nsfmax + 16);   // per atom
    b += nat * cap * (16 + 3 * R);                                                           // list slots
    b += nat * (4 * (size_t)triplet_capacity(entry_capacity(f.est, cap_hint_, 1024)) + 4 * (1 + kTripBlk));   // triplets, block tables
    b += (size_t)f.ncell * 12;                                                               // cells
    if ((size_t)f.ps * 2 <= kPosStartBytes) b += (size_t)f.ps * 2;                          // slot table
    b += (size_t)f.nstruct * (sizeof(StructGeo) + 80 * nm + 64 + 48);                       // per structure
    b += (nat / kReduceChunk + (size_t)f.nstruct + 1) * 88;                                 // reduction chunks
    if (!slack) return b;
    return b + b / 4 + (size_t(8) << 20);   // growth slack of each buffer, sort temporaries
}

void Engine::build(const std::vector<StructIn> &st, const int *Z_dev, const double *pos_dev, hipStream_t s,
                   double rc_list, bool launch)
{
    const int nstruct = (int)st.size();
    const double rc = rc_list;
    prep_ = Prepared{};
    prep_.st = st;
    ++build_gen_;
    trip_valid_ = false;
    tblk_valid_ = false;
    if (!(model_.maxrc > 0)) throw Error(3, "potential has no symmetry functions (cutoff 0)");
    if (dm_.fallback_el < 0) throw Error(3, elem_why_[0].empty() ? "no element of the potential can be evaluated"
                                                                 : elem_why_[0]);
    long nat_l = 0;
    for (const StructIn &in : st) {
        if (in.nat < 0) throw Error(1, "nat < 0");
        nat_l += in.nat;
    }
    if (nat_l >= (1L << 31)) throw Error(6, "more than 2^31 atoms in one evaluation");
    const int nat = (int)nat_l;
    last_nat_ = 0;
    last_total_ = 0;
    prep_.nat = nat;
    if (nat == 0) return;

    // per-structure grids, offsets and the neighbour-capacity estimate
    hgeo_.resize(nstruct);
    const bool fp64 = precision_ == 0;
    int lds_lim = 0;
    for (int c : kStage)
        if (sf_lds_bytes(fp64, c, dm_, true) <= 65536 && sf_lds_bytes(fp64, c, dm_, false) <= 65536)
            lds_lim = c;
    if (lds_lim == 0) throw Error(6, "symmetry-function tables too large for this build");
    long cell_off = 0, ps_total = 0;
    int atom_off = 0;
    double est = 0;
    for (int k = 0; k < nstruct; ++k) {
        StructGeo &G = hgeo_[k];
        structure_grid(st[k], rc, G);
        G.atom_off = atom_off;
        G.cell_off = (int)cell_off;
        G.ps_off = ps_total;
        atom_off += G.nat;
        cell_off += G.ncell;
        ps_total += (long)G.nat * G.S;
        if (cell_off > (1L << 28)) throw Error(6, "cell grids too large");
        est = std::max(est, neighbour_estimate(G, rc));
    }
    const int ncell = (int)cell_off;

    // per-atom buffers
    eidx_.ensure(4 * (size_t)nat);
    wpos_.ensure(sizeof(double4) * (size_t)nat);
    cellkey_.ensure(4 * (size_t)nat);
    cellkey_sorted_.ensure(4 * (size_t)nat);
    cell_atoms_.ensure(4 * (size_t)nat);
    iota_.ensure(4 * (size_t)nat);
    count_.ensure(4 * (size_t)nat);
    perm_.ensure(4 * (size_t)nat);
    Ei_.ensure(8 * (size_t)nat * dmm_.size());
    vir_.ensure(72 * (size_t)nat);
    sid_.ensure(4 * (size_t)nat);
    csid_.ensure(4 * (size_t)ncell);

    // reduction chunks: never across a structure
    std::vector<int2> chunks;
    std::vector<int> sc(nstruct + 1);
    for (int k = 0; k < nstruct; ++k) {
        sc[k] = (int)chunks.size();
        const StructGeo &G = hgeo_[k];
        for (int a = 0; a < G.nat; a += kReduceChunk)
            chunks.push_back(make_int2(G.atom_off + a, G.atom_off + std::min(G.nat, a + kReduceChunk)));
    }
    sc[nstruct] = (int)chunks.size();
    const int nchunk = (int)chunks.size();

    // host -> device tables through the pinned staging buffer
    const size_t geo_b = sizeof(StructGeo) * nstruct, ch_b = sizeof(int2) * nchunk, sc_b = 4 * (nstruct + 1);
    const int nm = (int)dmm_.size();
    const size_t res_b = 8 * 10 * (size_t)nstruct * nm;
    const size_t o_geo = 0, o_ch = (geo_b + 255) / 256 * 256, o_sc = o_ch + (ch_b + 255) / 256 * 256,
                 o_res = o_sc + (sc_b + 255) / 256 * 256, o_flags = o_res + (res_b + kFlagBytes + 255) / 256 * 256,
                 o_bbox = o_flags + 256;
    char *hs = (char *)staging(o_bbox + 256);
    // the tables are uploaded only when they differ from the last upload
    // (repeated evaluations of one geometry, as in MD, skip the copies)
    const bool same = uploaded_.size() == o_res && geo_.p && chunks_.p && sc_.p &&
                      std::memcmp(uploaded_.data(), hgeo_.data(), geo_b) == 0 &&
                      std::memcmp(uploaded_.data() + o_ch, chunks.data(), ch_b) == 0 &&
                      std::memcmp(uploaded_.data() + o_sc, sc.data(), sc_b) == 0;
    res_.ensure(res_b + kFlagBytes);   // the results, then a copy of the flags
    partial_.ensure(8 * 10 * (size_t)std::max(nchunk, 1));
    if (!same) {
        std::memcpy(hs + o_geo, hgeo_.data(), geo_b);
        std::memcpy(hs + o_ch, chunks.data(), ch_b);
        std::memcpy(hs + o_sc, sc.data(), sc_b);
        geo_.ensure(geo_b);
        chunks_.ensure(std::max<size_t>(ch_b, 8));
        sc_.ensure(sc_b);
        HIPCHK(hipMemcpyAsync(geo_.p, hs + o_geo, geo_b, hipMemcpyHostToDevice, s));
        if (nchunk) HIPCHK(hipMemcpyAsync(chunks_.p, hs + o_ch, ch_b, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(sc_.p, hs + o_sc, sc_b, hipMemcpyHostToDevice, s));
        uploaded_.assign(hs, hs + o_res);
    }

    int *mi = misc_.as<int>();
    Work w{};
    w.nat = nat;
    w.pos = pos_dev;
    w.Z = Z_dev;
    w.sid = sid_.as<int>();
    w.csid = csid_.as<int>();
    w.eidx = eidx_.as<int>();
    w.wpos = wpos_.as<double4>();
    w.cellkey = cellkey_.as<int>();
    w.cellkey_sorted = cellkey_sorted_.as<int>();
    w.cell_atoms = cell_atoms_.as<int>();
    w.count = count_.as<int>();
    w.perm = perm_.as<int>();
    w.elem_count = mi + kIntElemCount;
    w.elem_off = mi + kIntElemOff;
    w.flags = mi + kIntFlags;
    w.Ei = Ei_.as<double>();
    w.vir = vir_.as<double>();
    w.force = nullptr;              // set by finish
    w.result = res_.as<double>();

    Geo g{};
    g.sg = geo_.as<StructGeo>();
    g.rc2 = rc * rc;
    g.nstruct = nstruct;
    g.ncell = ncell;
    g.nat = nat;
    int bits = 1;
    while ((1L << bits) < ncell) ++bits;

    cell_start_.ensure(4 * (size_t)ncell);
    cell_end_.ensure(4 * (size_t)ncell);
    w.cell_start = cell_start_.as<int>();
    w.cell_end = cell_end_.as<int>();
    const bool small_sort = nat <= kSmallSortAtoms && ncell <= kSmallSortCells;
    // the counting sort ranks an atom by a scan of its cell (k^2 loads for a
    // cell of k atoms): dense grids keep the radix sort
    const bool count_sort = !small_sort && ncell <= kCountSortCells && (long)nat <= 64L * ncell;
    if (!small_sort) {   // the radix sort's temporaries and the element sort's chunks
        if (!count_sort) {
            size_t b1 = 0;
            HIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, b1, w.cellkey, w.cellkey_sorted, iota_.as<int>(),
                                                      w.cell_atoms, nat, 0, bits, s));
            cub_tmp_.ensure(b1);
        }
        elem_chunks_.ensure(4 * (size_t)kMaxEl * (size_t)std::max(1, elem_sort_chunks(nat)));
    }

    // Entry capacity per atom: the most neighbours seen by this context, or an
    // estimate from the mean density, rounded up to a staging size of the
    // symmetry-function kernels. An atom with more neighbours raises flags[1]
    // and the evaluation is repeated with room for it.
    w.lim = entry_capacity(est, cap_hint_, lds_lim);
    w.cap = slot_stride(w.lim);
    w.tcap = triplet_capacity(w.lim);
    w.trip = nullptr;               // set by compute when the passes share triplets
    w.tcount = nullptr;
    w.tblk = nullptr;
    if ((long)nat * w.cap >= (1L << 31)) throw Error(6, "neighbour list exceeds 2^31 entry slots");
    const size_t slots = (size_t)nat * w.cap;
    ent_j_.ensure(4 * slots);
    ent_shift_.ensure(4 * slots);
    ent_s_.ensure(4 * slots);
    rev_.ensure(4 * slots);
    w.ent_j = ent_j_.as<int>();
    w.ent_shift = ent_shift_.as<int>();
    w.ent_s = ent_s_.as<int>();
    w.rev = rev_.as<int>();
    w.pos_start = nullptr;
    if ((size_t)ps_total * 2 <= kPosStartBytes) {
        pos_start_.ensure((size_t)ps_total * 2);
        w.pos_start = pos_start_.as<unsigned short>();
    }
    prep_.w = w;
    prep_.g = g;
    prep_.nchunk = nchunk;
    prep_.lds_lim = lds_lim;
    prep_.ncell = ncell;
    prep_.bits = bits;
    prep_.small_sort = small_sort;
    prep_.count_sort = count_sort;
    prep_.o_res = o_res;
    prep_.o_flags = o_flags;
    prep_.res_b = res_b;
    prep_.valid = true;
    if (launch) build_launch(s, true);
}

void Engine::build_launch(hipStream_t s, bool timed)
{
    timed = timed && stage_timing_;
    ev_recorded_ = timed;
    if (prep_.nat == 0) return;
    if (!prep_.valid) throw Error(5, "no prepared evaluation");
    Work w = prep_.w;
    const Geo g = prep_.g;
    if (timed) HIPCHK(hipEventRecord(ev_[0], s));
    // flags and element counts start at zero (cleared by a kernel: no memset
    // node when the launches are captured)
    if (prep_.small_sort) {
        // one workgroup clears them, does the set-up and sorts by cell and by
        // element; the radix sort's input and output keys serve as scratch
        launch_sort_small(dm_, g, w, iota_.as<int>(), w.cellkey_sorted, misc_.as<int>(),
                          (int)(kDblOffset / sizeof(int)), s);
    } else if (prep_.count_sort) {
        // the flags, the cells' counts and the element sort's chunk counts
        // start at zero; the identity order's buffer holds the arrival slots,
        // the sorted keys' buffer the placement by slot
        int *ecc = elem_chunks_.as<int>();
        launch_zero3(misc_.as<int>(), (long)(kDblOffset / sizeof(int)), w.cell_end, prep_.ncell, ecc,
                     (long)kMaxEl * elem_sort_chunks(prep_.nat), s);
        launch_prep(dm_, g, w, nullptr, iota_.as<int>(), ecc, s);
        launch_cell_count_sort(w, prep_.ncell, dm_.nel, iota_.as<int>(), w.cellkey_sorted, ecc, s);
    } else {
        launch_fill_int(misc_.as<int>(), 0, (long)(kDblOffset / sizeof(int)), s);
        launch_prep(dm_, g, w, iota_.as<int>(), nullptr, nullptr, s);
        size_t bt = cub_tmp_.bytes;
        HIPCHK(hipcub::DeviceRadixSort::SortPairs(cub_tmp_.p, bt, w.cellkey, w.cellkey_sorted, iota_.as<int>(),
                                                  w.cell_atoms, prep_.nat, 0, prep_.bits, s));
        launch_cell_bounds(w, prep_.ncell, s);
        // atoms sorted by element (MLP tiles); needs no host round trip
        launch_elem_sort(w, dm_.nel, elem_chunks_.as<int>(), s);
    }
    HIPCHK(hipGetLastError());
    if (timed) HIPCHK(hipEventRecord(ev_[1], s));
    launch_nl_build(g, w, prep_.ncell, s);
    launch_nl_rev(g, w, s);
    HIPCHK(hipGetLastError());
    if (timed) HIPCHK(hipEventRecord(ev_[2], s));
}

bool Engine::finish(double *force_dev, hipStream_t s, double *energy, double *dEdLat)
{
    finish_launch(force_dev, s);
    return finish_collect(s, energy, dEdLat);
}

std::vector<char> Engine::launch_key() const
{
    const void *ptrs[] = {G_.p, dEdG_.p, gent_.p, chunks_.p, sc_.p, partial_.p, res_.p, misc_.p, trip_.p, tcount_.p};
    const int ints[] = {prep_.nat, prep_.nchunk, (int)prep_.st.size(), precision_, prep_.valid ? 1 : 0, tile_};
    std::vector<char> k(sizeof(Work) + sizeof(Geo) + sizeof ptrs + sizeof ints);
    char *p = k.data();
    std::memcpy(p, &prep_.w, sizeof(Work));
    p += sizeof(Work);
    std::memcpy(p, &prep_.g, sizeof(Geo));
    p += sizeof(Geo);
    std::memcpy(p, ptrs, sizeof ptrs);
    p += sizeof ptrs;
    std::memcpy(p, ints, sizeof ints);
    return k;
}

void Engine::finish_launch(double *force_dev, hipStream_t s, bool timed)
{
    timed = timed && stage_timing_;
    if (prep_.nat == 0) return;
    if (!prep_.valid) throw Error(5, "no prepared evaluation");
    const bool fp64 = precision_ == 0;
    Work w = prep_.w;
    w.force = force_dev;
    prep_.w.force = force_dev;
    const Geo g = prep_.g;
    const int nat = prep_.nat, nchunk = prep_.nchunk, nm = (int)dmm_.size();
    const int nstruct = (int)prep_.st.size();
    if (fp64)
        compute<double>(w, g, s, timed);
    else
        compute<float>(w, g, s, timed);
    for (int k = 0; k < nm; ++k) {       // per member: sum of its E_i (and the common virial)
        Work wk = w;
        wk.Ei = w.Ei + (size_t)k * nat;
        wk.result = w.result + 10 * (size_t)k * nstruct;
        int *flag_out = k == 0 ? (int *)(w.result + 10 * (size_t)nm * nstruct) : nullptr;
        // the first member's reduction also adds the partnerless entries
        unsigned *done = (unsigned *)(misc_.as<int>() + kIntReduceDone);
        if (fp64)
            launch_reduce<double>(wk, chunks_.as<int2>(), nchunk, sc_.as<int>(), nstruct, partial_.as<double>(),
                                  flag_out, done, k == 0 ? gent_.as<double>() : nullptr, s);
        else
            launch_reduce<float>(wk, chunks_.as<int2>(), nchunk, sc_.as<int>(), nstruct, partial_.as<double>(),
                                 flag_out, done, k == 0 ? gent_.as<float>() : nullptr, s);
    }
    HIPCHK(hipGetLastError());
    if (timed) HIPCHK(hipEventRecord(ev_[6], s));
}

bool Engine::finish_collect(hipStream_t s, double *energy, double *dEdLat)
{
    const int nstruct = (int)prep_.st.size();
    for (int k = 0; k < nstruct; ++k) energy[k] = 0;
    if (dEdLat) std::memset(dEdLat, 0, 9 * sizeof(double) * nstruct);
    if (prep_.nat == 0) return true;
    if (!prep_.valid) throw Error(5, "no prepared evaluation");
    const Work &w = prep_.w;
    const int nat = prep_.nat, nm = (int)dmm_.size();
    const std::vector<StructIn> &st = prep_.st;
    char *hs = (char *)stage_ptr_;   // where build staged its tables
    double *hres = (double *)(hs + prep_.o_res);
    int *hflags = (int *)(hs + prep_.o_res + prep_.res_b);   // the reduction's copy of w.flags
    HIPCHK(hipMemcpyAsync(hres, w.result, prep_.res_b + kFlagBytes, hipMemcpyDeviceToHost, s));
    if (pending_force_.bytes)
        HIPCHK(hipMemcpyAsync(pending_force_.host, pending_force_.dev, pending_force_.bytes, hipMemcpyDeviceToHost,
                              s));
    HIPCHK(hipStreamSynchronize(s));
    if (hflags[0]) throw Error(4, "structure contains element Z=" + std::to_string(hflags[0]) +
                                      " that the potential does not describe");
    if (hflags[3]) {
        for (int e = 0; e < dm_.nel; ++e)
            if (model_.el[e].Z == hflags[3]) throw Error(3, elem_why_[e]);
        throw Error(3, "element Z=" + std::to_string(hflags[3]) + " cannot be evaluated");
    }
    if (hflags[1] > 0) {
        if (hflags[1] > prep_.lds_lim)
            throw Error(6, "an atom has " + std::to_string(hflags[1]) + " neighbours; this build stages at most " +
                               std::to_string(prep_.lds_lim) + " per atom in LDS");
        cap_hint_ = std::max(cap_hint_, hflags[1]);
        prep_.valid = false;
        return false;
    }
    if (hflags[2]) throw Error(6, "periodic image shift beyond +-127 cells");
    last_nat_ = nat;
    last_cap_ = w.cap;
    last_lim_ = w.lim;
    last_total_ = -1;
    for (int k = 0; k < 6; ++k) {
        float ms = 0;
        if (ev_recorded_) HIPCHK(hipEventElapsedTime(&ms, ev_[k], ev_[k + 1]));
        ms_[k] = ms;
    }
    // committee mean and population standard deviation of the energies
    // (globalHdnnps.f90:57-67)
    member_E_.assign((size_t)nstruct * nm, 0.0);
    std_E_.assign(nstruct, 0.0);
    for (int k = 0; k < nstruct; ++k) {
        double mean = 0;
        for (int j = 0; j < nm; ++j) {
            member_E_[(size_t)k * nm + j] = hres[10 * ((size_t)j * nstruct + k)];
            mean += member_E_[(size_t)k * nm + j];
        }
        mean /= nm;
        double var = 0;
        for (int j = 0; j < nm; ++j) var += (member_E_[(size_t)k * nm + j] - mean) * (member_E_[(size_t)k * nm + j] - mean);
        std_E_[k] = std::sqrt(var / nm);
        energy[k] = nm == 1 ? hres[10 * (size_t)k] : mean;
    }
    for (int k = 0; k < nstruct; ++k) {
        const double *res = hres + 10 * (size_t)k;
        if (dEdLat && st[k].periodic) {   // dEdLat(a,x) = sum_b W(a,b) inv(lat)(x,b), stored at [3x + a]
            double il[9];
            inv3(st[k].lat, il);
            double *dl = dEdLat + 9 * (size_t)k;
            for (int a = 0; a < 3; ++a)
                for (int x = 0; x < 3; ++x) {
                    double acc = 0;
                    for (int b = 0; b < 3; ++b) acc += res[1 + 3 * a + b] * il[x + 3 * b];
                    dl[3 * x + a] = acc;
                }
        }
    }
    return true;
}

void Engine::run(std::vector<StructIn> &st, const int *Z_dev, const double *pos_dev, double *force_dev,
                 hipStream_t s, double *energy, double *dEdLat, int retry)
{
    const double rc = model_.maxrc;
    for (;;) {
        build(st, Z_dev, pos_dev, s, rc, false);
        launch_evaluation(force_dev, s);
        if (finish_collect(s, energy, dEdLat)) return;
        if (retry-- <= 0) throw Error(6, "neighbour capacity did not converge");
    }
}

// Everything the kernels of build_launch and finish_launch take (launch
// parameters, buffer addresses, the device models): a graph captured from
// them replays the same launches while this stays the same.
std::vector<char> Engine::eval_key() const
{
    std::vector<char> k = launch_key();
    auto add = [&k](const void *p, size_t n) {
        const char *c = (const char *)p;
        k.insert(k.end(), c, c + n);
    };
    const int ints[] = {prep_.ncell, prep_.bits, prep_.small_sort ? 1 : 0, skip_values_ ? 1 : 0,
                        (int)dmm_.size(), mlp_maxw_, mlp_sumc_};
    add(ints, sizeof ints);
    for (const DevModel &d : dmm_) add(&d, sizeof d);
    add(&dm_, sizeof dm_);
    add(maxw_m_.data(), sizeof(int) * maxw_m_.size());
    std::vector<const void *> ptrs;
    for (DevMem *b : const_cast<Engine *>(this)->work_buffers()) ptrs.push_back(b->p);
    ptrs.push_back(misc_.p);
    ptrs.push_back(stream_);
    add(ptrs.data(), sizeof(void *) * ptrs.size());
    return k;
}

void Engine::launch_evaluation(double *force_dev, hipStream_t s)
{
    const bool graphable = eval_graphs_on_ && prep_.valid && prep_.nat > 0 && prep_.small_sort && s == stream_ &&
                           !t_arena && !skip_values_;
    if (!graphable) {
        build_launch(s, true);
        finish_launch(force_dev, s, true);
        return;
    }
    prep_.w.force = force_dev;
    std::vector<char> key = eval_key();
    EvalGraph *hit = nullptr;
    for (EvalGraph &e : eval_graphs_)
        if (e.key == key) hit = &e;
    if (!hit && key == eval_seen_) hit = capture_evaluation(force_dev, s, key);
    if (!hit) {
        build_launch(s, true);
        finish_launch(force_dev, s, true);
        eval_seen_ = eval_key();   // with the buffers this launch sized
        return;
    }
    hit->used = ++eval_tick_;
    // the replay's whole device time is the first stage's
    ev_recorded_ = stage_timing_;
    if (ev_recorded_) HIPCHK(hipEventRecord(ev_[0], s));
    HIPCHK(hipGraphLaunch(hit->exec, s));
    if (ev_recorded_)
        for (int k = 1; k < 7; ++k) HIPCHK(hipEventRecord(ev_[k], s));
    ++eval_replays_;
}

Engine::EvalGraph *Engine::capture_evaluation(double *force_dev, hipStream_t s, std::vector<char> key)
{
    // no other thread allocates, frees or copies synchronously through this
    // library while the capture runs (HipUnsafe)
    HipUnsafe::Capture excl;
    if (hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed) != hipSuccess) {
        (void)hipGetLastError();
        eval_graphs_on_ = false;
        return nullptr;
    }
    bool ok = true;
    try {
        build_launch(s, false);
        finish_launch(force_dev, s, false);
    } catch (const std::exception &) {
        ok = false;
    }
    hipGraph_t g = nullptr;
    ok = hipStreamEndCapture(s, &g) == hipSuccess && ok && g;
    hipGraphExec_t x = nullptr;
    if (ok) ok = hipGraphInstantiate(&x, g, nullptr, nullptr, 0) == hipSuccess;
    if (!ok) {   // launch directly from now on
        if (x) (void)hipGraphExecDestroy(x);
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        eval_graphs_on_ = false;
        return nullptr;
    }
    if ((int)eval_graphs_.size() >= kEvalGraphs) {   // the least recently used goes
        size_t lru = 0;
        for (size_t k = 1; k < eval_graphs_.size(); ++k)
            if (eval_graphs_[k].used < eval_graphs_[lru].used) lru = k;
        (void)hipGraphExecDestroy(eval_graphs_[lru].exec);
        (void)hipGraphDestroy(eval_graphs_[lru].graph);
        eval_graphs_.erase(eval_graphs_.begin() + lru);
    }
    eval_graphs_.push_back(EvalGraph{std::move(key), g, x, ++eval_tick_});
    ++eval_captures_;
    return &eval_graphs_.back();
}

void Engine::drop_eval_graphs()
{
    for (EvalGraph &e : eval_graphs_) {
        (void)hipGraphExecDestroy(e.exec);
        (void)hipGraphDestroy(e.graph);
    }
    eval_graphs_.clear();
    eval_seen_.clear();
}

void Engine::set_eval_graphs(bool on)
{
    eval_graphs_on_ = on;
    if (!on) drop_eval_graphs();
}

void Engine::eval_graph_counts(int64_t *captures, int64_t *replays) const
{
    if (captures) *captures = eval_captures_;
    if (replays) *replays = eval_replays_;
}

void Engine::evaluate_batch_host(int nstruct, const int *nat, const int *Z, const double *pos, const int *periodic,
                                 const double *lat, double *energy, double *force, double *dEdLat)
{
    if (!dev_ready_) init_device();
    if (nstruct < 0) throw Error(1, "nstruct < 0");
    if (nstruct == 0) return;
    if (!nat || !periodic || !energy) throw Error(1, "null array");
    std::vector<StructIn> st(nstruct);
    long total = 0;
    for (int k = 0; k < nstruct; ++k) {
        StructIn &in = st[k];
        in.nat = nat[k];
        in.periodic = periodic[k] ? 1 : 0;
        if (in.nat < 0) throw Error(1, "nat < 0");
        if (in.periodic) {
            if (!lat) throw Error(1, "periodic structure without lattice");
            std::memcpy(in.lat, lat + 9 * (size_t)k, sizeof in.lat);
        } else if (in.nat > 0) {           // bounding box of the raw positions (neighborLists.f90:131-138)
            const double *p = pos + 3 * (size_t)total;
            for (int a = 0; a < 3; ++a) in.bbox[a] = in.bbox[3 + a] = p[a];
            for (int i = 1; i < in.nat; ++i)
                for (int a = 0; a < 3; ++a) {
                    in.bbox[a] = std::min(in.bbox[a], p[3 * i + a]);
                    in.bbox[3 + a] = std::max(in.bbox[3 + a], p[3 * i + a]);
                }
        }
        total += in.nat;
    }
    if (total > 0 && (!Z || !pos || !force)) throw Error(1, "null array");
    if (total >= (1L << 31)) throw Error(6, "more than 2^31 atoms in one evaluation");
    const size_t n = (size_t)std::max<long>(total, 1);
    const size_t pos_b = 24 * (size_t)total, z_off = (24 * n + 255) / 256 * 256, in_b = z_off + 4 * (size_t)total;
    posin_.ensure(z_off + 4 * n);
    fout_.ensure(24 * n);
    const int *Z_dev = (const int *)(posin_.as<char>() + z_off);
    // Small inputs go through pinned staging: one upload, and the forces come
    // back with the results before the one synchronisation. Large ones are
    // copied straight from the caller's arrays (a staging copy would cost more
    // than the launches it saves).
    const bool small = total > 0 && total <= kSmallHostEval;
    char *io = small ? (char *)io_staging(std::max(in_b, pos_b)) : nullptr;
    if (small) {
        std::memcpy(io, pos, pos_b);
        std::memcpy(io + z_off, Z, 4 * (size_t)total);
        HIPCHK(hipMemcpyAsync(posin_.p, io, in_b, hipMemcpyHostToDevice, stream_));
        pending_force_ = {(double *)io, fout_.as<double>(), pos_b};
    } else if (total > 0) {
        HIPCHK(hipMemcpyAsync((void *)Z_dev, Z, 4 * (size_t)total, hipMemcpyHostToDevice, stream_));
        HIPCHK(hipMemcpyAsync(posin_.p, pos, pos_b, hipMemcpyHostToDevice, stream_));
    }
    try {
        run(st, Z_dev, posin_.as<double>(), fout_.as<double>(), stream_, energy, dEdLat, 1);
    } catch (...) {
        pending_force_ = {};
        throw;
    }
    pending_force_ = {};
    if (small) {
        std::memcpy(force, io, pos_b);
    } else if (total > 0) {
        HIPCHK(hipMemcpyAsync(force, fout_.p, pos_b, hipMemcpyDeviceToHost, stream_));
        HIPCHK(hipStreamSynchronize(stream_));
    }
}

void Engine::set_tile(int atoms_per_wave)
{
    if (atoms_per_wave != 1 && atoms_per_wave != 2 && atoms_per_wave != 4)
        throw Error(1, "the tile (atoms per wavefront) must be 1, 2 or 4");
    tile_ = atoms_per_wave;
}

int Engine::autotune(int nstruct, const int *nat, const int *Z, const double *pos, const int *periodic,
                     const double *lat, double *ms)
{
    if (nstruct < 1 || !nat) throw Error(1, "autotune needs a sample of at least one structure");
    long total = 0;
    for (int k = 0; k < nstruct; ++k) total += std::max(0, nat[k]);
    std::vector<double> E(nstruct), F(3 * (size_t)std::max<long>(total, 1));
    const int cand[3] = {1, 2, 4};
    const int keep = tile_;
    const bool graphs = eval_graphs_on_, timing = stage_timing_;
    eval_graphs_on_ = false;   // the stage timings of direct launches
    stage_timing_ = true;
    int best = 1;
    double best_ms = 0;
    try {
        for (int c = 0; c < 3; ++c) {
            tile_ = cand[c];
            evaluate_batch_host(nstruct, nat, Z, pos, periodic, lat, E.data(), F.data(), nullptr);   // warm-up
            double t = 0;
            if (sf_tile(precision_ == 0, last_lim_, cand[c], dm_.nsfmax, dm_.g4_max, dm_.nel) == cand[c]) {
                for (int rep = 0; rep < 3; ++rep) {
                    evaluate_batch_host(nstruct, nat, Z, pos, periodic, lat, E.data(), F.data(), nullptr);
                    const double v = ms_[2] + ms_[4];
                    t = rep == 0 ? v : std::min(t, v);
                }
                if (c == 0 || t < best_ms) {
                    best = cand[c];
                    best_ms = t;
                }
            }
            if (ms) ms[c] = t;
        }
    } catch (...) {
        tile_ = keep;
        eval_graphs_on_ = graphs;
        stage_timing_ = timing;
        throw;
    }
    tile_ = best;
    eval_graphs_on_ = graphs;
    stage_timing_ = timing;
    return best;
}

void Engine::evaluate_host(int nat, const int *Z, const double *pos, int periodic, const double *lat,
                           double *energy, double *force, double *dEdLat)
{
    if (nat > 0 && (!Z || !pos || !force)) throw Error(1, "null array");
    evaluate_batch_host(1, &nat, Z, pos, &periodic, lat, energy, force, dEdLat);
}

void Engine::evaluate_batch_device(int nstruct, const int *nat, const int *Z_dev, const double *pos_dev,
                                   const int *periodic, const double *lat, double *energy, double *force_dev,
                                   double *dEdLat, hipStream_t stream)
{
    if (!dev_ready_) init_device();
    if (nstruct < 0) throw Error(1, "nstruct < 0");
    if (nstruct == 0) return;
    if (!nat || !periodic || !energy) throw Error(1, "null array");
    hipStream_t s = stream ? stream : stream_;
    std::vector<StructIn> st(nstruct);
    long total = 0;
    for (int k = 0; k < nstruct; ++k) {
        StructIn &in = st[k];
        in.nat = nat[k];
        in.periodic = periodic[k] ? 1 : 0;
        if (in.nat < 0) throw Error(1, "nat < 0");
        total += in.nat;
        if (total >= (1L << 31)) throw Error(6, "more than 2^31 atoms in one evaluation");
        if (in.periodic) {
            if (!lat) throw Error(1, "periodic structure without lattice");
            std::memcpy(in.lat, lat + 9 * (size_t)k, sizeof in.lat);
        }
    }
    if (total > 0 && (!Z_dev || !pos_dev || !force_dev)) throw Error(1, "null array");
    device_bboxes(st, pos_dev, s);
    if (t_arena) {   // refuse what cannot fit before anything is launched
        Footprint f;
        for (const StructIn &in : st) footprint_add(f, in, model_.maxrc);
        const size_t need = footprint_bytes(f, false);
        if (need > t_arena->size - t_arena->used)
            throw Error(6, "the evaluation needs about " + std::to_string((need + (1 << 20) - 1) >> 20) +
                               " MiB of device scratch; the entry point that allocates nothing has " +
                               std::to_string(t_arena->size >> 20) + " MiB");
    }
    run(st, Z_dev, pos_dev, force_dev, s, energy, dEdLat, 1);
}

void Engine::evaluate_batch_device_noalloc(int nstruct, const int *nat, const int *Z_dev, const double *pos_dev,
                                           const int *periodic, const double *lat, double *energy, double *force_dev,
                                           double *dEdLat, hipStream_t stream)
{
    if (!dev_ready_)
        throw Error(5, "the context has no device state (no GPU was visible when it was created), and the "
                       "entry point that allocates nothing cannot set it up");
    int cur = -1;
    HIPCHK(hipGetDevice(&cur));
    if (cur != dev_) throw Error(1, "the current device is not the one the context was set up on");
    // one call at a time per device: they share its scratch
    static std::mutex dev_mu[64];
    std::lock_guard<std::mutex> one(dev_mu[dev_ % 64]);
    // the context's own buffers are set aside and put back afterwards
    std::vector<DevMem *> bufs = work_buffers();
    bufs.push_back(&misc_);
    std::vector<DevMem> saved(bufs.size());
    for (size_t k = 0; k < bufs.size(); ++k) {
        saved[k] = *bufs[k];
        *bufs[k] = DevMem{};
    }
    Arena arena{(char *)scratch_, kScratchBytes, 0};
    auto restore = [&] {
        t_arena = nullptr;
        for (size_t k = 0; k < bufs.size(); ++k) *bufs[k] = saved[k];
        uploaded_.clear();   // the tables were staged into the scratch
        prep_ = Prepared{};
        ++build_gen_;        // lists built from the scratch are gone
        last_nat_ = 0;
        last_total_ = 0;
    };
    t_arena = &arena;
    try {
        uploaded_.clear();
        misc_.ensure(1024);
        evaluate_batch_device(nstruct, nat, Z_dev, pos_dev, periodic, lat, energy, force_dev, dEdLat, stream);
    } catch (...) {
        (void)hipStreamSynchronize(stream ? stream : stream_);   // nothing left running on the scratch
        restore();
        throw;
    }
    restore();
}

// bounding boxes of the nonperiodic structures of st (atoms concatenated in
// pos_dev), computed on the device (neighborLists.f90:131-138)
void Engine::device_bboxes(std::vector<StructIn> &st, const double *pos_dev, hipStream_t s)
{
    const int nstruct = (int)st.size();
    std::vector<int> off(nstruct + 1, 0), which;
    for (int k = 0; k < nstruct; ++k) {
        off[k + 1] = off[k] + st[k].nat;
        if (!st[k].periodic && st[k].nat > 0) which.push_back(k);
    }
    if (!which.empty()) {
        const int nw = (int)which.size();
        const size_t ob = 4 * (size_t)(nstruct + 1), wb = 4 * (size_t)nw, bb = 48 * (size_t)nw;
        char *hs = (char *)staging(ob + wb + bb + 512);
        std::memcpy(hs, off.data(), ob);
        std::memcpy(hs + (ob + 255) / 256 * 256, which.data(), wb);
        double *hb = (double *)(hs + (ob + 255) / 256 * 256 + (wb + 255) / 256 * 256);
        bbox_.ensure(ob + wb + bb + 512);
        char *db = (char *)bbox_.p;
        HIPCHK(hipMemcpyAsync(db, hs, (char *)hb - hs, hipMemcpyHostToDevice, s));
        launch_bbox(pos_dev, (const int *)db, (const int *)(db + ((ob + 255) / 256 * 256)), nw,
                    (double *)(db + ((char *)hb - hs)), s);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(hb, db + ((char *)hb - hs), bb, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        for (int t = 0; t < nw; ++t) std::memcpy(st[which[t]].bbox, hb + 6 * (size_t)t, 6 * sizeof(double));
    }
}

void Engine::evaluate_device(int nat, const int *Z_dev, const double *pos_dev, int periodic,
                             const double *lat, double *energy, double *force_dev, double *dEdLat,
                             hipStream_t stream)
{
    if (periodic && !lat) throw Error(1, "periodic structure without lattice");
    evaluate_batch_device(1, &nat, Z_dev, pos_dev, &periodic, lat, energy, force_dev, dEdLat, stream);
}

// ---------------------------------------------------------------- stage API
// hdnnpCalcNeighborLists / hdnnpCalcSymfunctions / hdnnpCalcNNs
// (hdnnps.f90:373-562) on the engine's one stage.

int64_t Engine::stage_build(int nat, const int *Z, const double *pos, int periodic, const double *lat)
{
    if (!dev_ready_) init_device();
    if (nat < 0) throw Error(1, "nat < 0");
    if (nat > 0 && (!Z || !pos)) throw Error(1, "null array");
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
    const size_t n = (size_t)std::max(nat, 1), z_off = (24 * n + 255) / 256 * 256;
    posin_.ensure(z_off + 4 * n);
    fout_.ensure(24 * n);
    const int *Z_dev = (const int *)(posin_.as<char>() + z_off);
    if (nat > 0) {
        HIPCHK(hipMemcpyAsync(posin_.p, pos, 24 * (size_t)nat, hipMemcpyHostToDevice, stream_));
        HIPCHK(hipMemcpyAsync((void *)Z_dev, Z, 4 * (size_t)nat, hipMemcpyHostToDevice, stream_));
    }
    std::vector<StructIn> st(1, in);
    for (int attempt = 0;; ++attempt) {
        build(st, Z_dev, posin_.as<double>(), stream_, model_.maxrc);
        if (nat == 0) break;
        int fl[8];
        d2h(fl, misc_.as<int>() + kIntFlags, sizeof fl);
        if (fl[0]) throw Error(4, "structure contains element Z=" + std::to_string(fl[0]) +
                                      " that the potential does not describe");
        if (fl[3]) {
            for (int e = 0; e < dm_.nel; ++e)
                if (model_.el[e].Z == fl[3]) throw Error(3, elem_why_[e]);
            throw Error(3, "element Z=" + std::to_string(fl[3]) + " cannot be evaluated");
        }
        if (fl[2]) throw Error(6, "periodic image shift beyond +-127 cells");
        if (fl[1] <= 0) break;
        if (fl[1] > prep_.lds_lim)
            throw Error(6, "an atom has " + std::to_string(fl[1]) + " neighbours; this build stages at most " +
                               std::to_string(prep_.lds_lim) + " per atom in LDS");
        if (attempt >= 1) throw Error(6, "neighbour capacity did not converge");
        cap_hint_ = std::max(cap_hint_, fl[1]);
    }
    stage_in_ = in;
    stage_pos_.assign(pos, pos + 3 * (size_t)nat);
    stage_gen_ = build_gen_;
    stage_id_ = ++stage_next_;
    last_nat_ = nat;
    last_cap_ = prep_.w.cap;
    last_lim_ = prep_.w.lim;
    last_total_ = -1;
    return stage_id_;
}

bool Engine::stage_current(int64_t id) const
{
    return id != 0 && id == stage_id_ && build_gen_ == stage_gen_ && (prep_.valid || prep_.nat == 0);
}

void Engine::stage_neighbors(int32_t *nneis, int32_t *ineis, double *neis)
{
    if (!stage_current(stage_id_)) throw Error(1, "no current neighbour-list stage");
    const int nat = prep_.nat;
    if (nat == 0) return;
    const Work &w = prep_.w;
    std::vector<int> cnt(nat);
    d2h(cnt.data(), w.count, 4 * (size_t)nat);
    for (int i = 0; i < nat; ++i) nneis[i] = 1 + cnt[i];
    if (!ineis && !neis) return;
    const size_t slots = (size_t)nat * w.cap;
    std::vector<int> j(slots), sh(slots);
    std::vector<double4> wp(nat);
    d2h(j.data(), w.ent_j, 4 * slots);
    d2h(sh.data(), w.ent_shift, 4 * slots);
    d2h(wp.data(), w.wpos, sizeof(double4) * (size_t)nat);
    const StructGeo &G = hgeo_[0];
    // slot 1: the position the list was built from, once wrapped into the
    // cell when periodic (hdnnps.f90:380-382), as given otherwise; the
    // images: the list-frame position plus the lattice offset summed as
    // neighborLists.f90:186-209 does (z, y, x, wrapped axes only), plus the
    // artificial origin (:212)
    size_t q = 0;
    for (int i = 0; i < nat; ++i) {
        double p[3] = {stage_pos_[3 * (size_t)i], stage_pos_[3 * (size_t)i + 1], stage_pos_[3 * (size_t)i + 2]};
        if (G.periodic) {   // moveAtomsIntoCell (pointsets.f90:201-217) in the kernels' operation order
            double s3[3];
            for (int a = 0; a < 3; ++a) s3[a] = (G.inv[a] * p[0] + G.inv[3 + a] * p[1]) + G.inv[6 + a] * p[2];
            for (int a = 0; a < 3; ++a) s3[a] = s3[a] - std::floor(s3[a]);
            for (int a = 0; a < 3; ++a) p[a] = (G.lat[a] * s3[0] + G.lat[3 + a] * s3[1]) + G.lat[6 + a] * s3[2];
        }
        if (ineis) ineis[q] = i + 1;
        if (neis) std::memcpy(neis + 3 * q, p, sizeof p);
        ++q;
        for (int n = 0; n < cnt[i]; ++n, ++q) {
            const size_t t = (size_t)i * w.cap + n;
            if (ineis) ineis[q] = j[t] + 1;
            if (!neis) continue;
            int sx, sy, sz;
            unpack_shift(sh[t], sx, sy, sz);
            double dl[3] = {0, 0, 0};
            for (int a = 0; a < 3; ++a) {
                if (sz) dl[a] = dl[a] + G.lat[6 + a] * (double)sz;
                if (sy) dl[a] = dl[a] + G.lat[3 + a] * (double)sy;
                if (sx) dl[a] = dl[a] + G.lat[a] * (double)sx;
            }
            const double *pj = &wp[j[t]].x;
            for (int a = 0; a < 3; ++a) neis[3 * q + a] = (pj[a] + dl[a]) + G.origin[a];
        }
    }
}

void Engine::stage_symfunctions(int together, int ld, double *G, double *dsf)
{
    if (!stage_current(stage_id_)) throw Error(1, "no current neighbour-list stage");
    const int nat = prep_.nat;
    if (nat == 0) return;
    if (ld < dm_.nsfmax) throw Error(1, "ld smaller than the largest SF count");
    const bool fp64 = precision_ == 0;
    Work w = prep_.w;
    const Geo g = prep_.g;
    hipStream_t s = stream_;
    if (together) {   // the fused values pass
        const size_t rb = fp64 ? 8 : 4;
        G_.ensure(rb * (size_t)nat * dm_.nsfmax);
        const int ga = sf_tile(fp64, w.lim, tile_, dm_.nsfmax, dm_.g4_max, dm_.nel);
        w.trip = nullptr;
        w.tcount = nullptr;
        w.tblk = nullptr;
        if (ga == 1) {
            trip_.ensure(4 * (size_t)nat * w.tcap);
            tcount_.ensure(4 * (size_t)nat * (1 + kTripBlk));
            w.trip = trip_.as<int>();
            w.tcount = tcount_.as<int>();
            w.tblk = w.tcount + nat;
        }
        if (fp64) {
            launch_sf_values<double>(dm_, g, w, w.lim, ga, G_.as<double>(), s);
            if (dm_.g5_total) launch_g5_values<double>(dm_, g, w, G_.as<double>(), s);
        } else {
            launch_sf_values<float>(dm_, g, w, w.lim, ga, G_.as<float>(), s);
            if (dm_.g5_total) launch_g5_values<float>(dm_, g, w, G_.as<float>(), s);
        }
        HIPCHK(hipGetLastError());
        trip_valid_ = w.trip != nullptr;
        tblk_valid_ = w.tblk != nullptr;
        std::vector<double> tmp((size_t)nat * dm_.nsfmax);
        if (fp64) {
            d2h(tmp.data(), G_.p, 8 * tmp.size());
        } else {
            std::vector<float> t(tmp.size());
            d2h(t.data(), G_.p, 4 * t.size());
            for (size_t k = 0; k < t.size(); ++k) tmp[k] = t[k];
        }
        for (int i = 0; i < nat; ++i)
            for (int k = 0; k < ld; ++k) G[(size_t)i * ld + k] = k < dm_.nsfmax ? tmp[(size_t)i * dm_.nsfmax + k] : 0.0;
    }
    if (together && !dsf) return;
    // the per-function path: its parameter table, then per-atom row offsets
    if (!sfpar_.p) {
        std::vector<SfParam> par;
        std::vector<int> off(dm_.nel + 1, 0);
        for (int e = 0; e < dm_.nel; ++e) {
            off[e] = (int)par.size();
            for (const SymFn &f : model_.el[e].sfs) {
                SfParam p{};
                p.type = f.type;
                p.s = (int)par.size() - off[e];
                p.f = model_.rank_of_Z(f.e1);
                p.g = f.type != 2 ? model_.rank_of_Z(f.e2) : -1;
                p.eta = f.eta;
                p.rs = f.type == 2 ? f.rs : 0.0;
                p.lambda = f.type != 2 ? f.lambda : 0.0;
                p.zeta = f.type != 2 ? f.zeta : 1.0;
                p.pre = f.type != 2 ? std::pow(2.0, 1.0 - f.zeta) : 1.0;
                par.push_back(p);
            }
        }
        off[dm_.nel] = (int)par.size();
        sfpar_.ensure(sizeof(SfParam) * std::max<size_t>(1, par.size()));
        sfpar_off_.ensure(4 * off.size());
        HIPCHK(hipMemcpyAsync(sfpar_.p, par.data(), sizeof(SfParam) * par.size(), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(sfpar_off_.p, off.data(), 4 * off.size(), hipMemcpyHostToDevice, s));
        HIPCHK(hipStreamSynchronize(s));
    }
    std::vector<int> cnt(nat), el(nat);
    d2h(cnt.data(), w.count, 4 * (size_t)nat);
    d2h(el.data(), w.eidx, 4 * (size_t)nat);
    std::vector<long long> off(nat + 1, 0);
    for (int i = 0; i < nat; ++i) off[i + 1] = off[i] + 3LL * dm_.nsf[el[i]] * (cnt[i] + 1);
    stage_off_.ensure(8 * off.size());
    stage_G_.ensure(8 * (size_t)nat * ld);
    if (dsf) stage_dsf_.ensure(8 * (size_t)std::max<long long>(off[nat], 1));
    HIPCHK(hipMemcpyAsync(stage_off_.p, off.data(), 8 * off.size(), hipMemcpyHostToDevice, s));
    if (fp64)
        launch_sf_single<double>(dm_, g, w, sfpar_.as<SfParam>(), sfpar_off_.as<int>(), stage_off_.as<long long>(), ld,
                                 stage_G_.as<double>(), dsf ? stage_dsf_.as<double>() : nullptr, s);
    else
        launch_sf_single<float>(dm_, g, w, sfpar_.as<SfParam>(), sfpar_off_.as<int>(), stage_off_.as<long long>(), ld,
                                stage_G_.as<double>(), dsf ? stage_dsf_.as<double>() : nullptr, s);
    HIPCHK(hipGetLastError());
    if (!together) {
        std::vector<double> tmp((size_t)nat * ld);
        d2h(tmp.data(), stage_G_.p, 8 * tmp.size());
        for (int i = 0; i < nat; ++i)   // the columns an element does not have stay 0
            for (int k = 0; k < ld; ++k) G[(size_t)i * ld + k] = k < dm_.nsf[el[i]] ? tmp[(size_t)i * ld + k] : 0.0;
    }
    if (dsf) d2h(dsf, stage_dsf_.p, 8 * (size_t)off[nat]);
}

void Engine::stage_nns(Engine &net, int ld, const double *G, double *energy, double *force, double *dEdLat)
{
    if (!stage_current(stage_id_)) throw Error(1, "no current neighbour-list stage");
    if (net.precision_ != precision_) throw Error(1, "the networks' context has another precision");
    if (net.committee_size() != 1) throw Error(1, "hdnnpCalcNNs takes the networks of one potential");
    if (net.model_.el.size() != model_.el.size()) throw Error(3, "the potentials describe different elements");
    for (size_t e = 0; e < model_.el.size(); ++e)
        if (net.model_.el[e].Z != model_.el[e].Z || net.model_.el[e].sfs.size() != model_.el[e].sfs.size())
            throw Error(3, "the networks' potential does not take this potential's symmetry functions");
    net.ensure_device();
    const int nat = prep_.nat;
    *energy = 0;
    if (dEdLat) std::memset(dEdLat, 0, 9 * sizeof(double));
    if (nat == 0) return;
    if (ld < dm_.nsfmax) throw Error(1, "ld smaller than the largest SF count");
    const bool fp64 = precision_ == 0;
    G_.ensure((fp64 ? 8 : 4) * (size_t)nat * dm_.nsfmax);
    if (fp64) {
        std::vector<double> t((size_t)nat * dm_.nsfmax);
        for (int i = 0; i < nat; ++i)
            for (int k = 0; k < dm_.nsfmax; ++k) t[(size_t)i * dm_.nsfmax + k] = G[(size_t)i * ld + k];
        HIPCHK(hipMemcpyAsync(G_.p, t.data(), 8 * t.size(), hipMemcpyHostToDevice, stream_));
        HIPCHK(hipStreamSynchronize(stream_));
    } else {
        std::vector<float> t((size_t)nat * dm_.nsfmax);
        for (int i = 0; i < nat; ++i)
            for (int k = 0; k < dm_.nsfmax; ++k) t[(size_t)i * dm_.nsfmax + k] = (float)G[(size_t)i * ld + k];
        HIPCHK(hipMemcpyAsync(G_.p, t.data(), 4 * t.size(), hipMemcpyHostToDevice, stream_));
        HIPCHK(hipStreamSynchronize(stream_));
    }
    // the networks of `net` on this engine's symmetry functions and lists
    std::vector<DevModel> keep_m(1, net.dm_);
    std::vector<int> keep_w(1, net.mlp_maxw_);
    std::swap(dmm_, keep_m);
    std::swap(maxw_m_, keep_w);
    skip_values_ = true;
    auto restore = [&] {
        std::swap(dmm_, keep_m);
        std::swap(maxw_m_, keep_w);
        skip_values_ = false;
    };
    double dl[9];
    try {
        if (!finish(fout_.as<double>(), stream_, energy, dl)) throw Error(6, "neighbour capacity exceeded");
    } catch (...) {
        restore();
        throw;
    }
    restore();
    if (dEdLat) std::memcpy(dEdLat, dl, sizeof dl);
    d2h(force, fout_.p, 24 * (size_t)nat);
}

// a device-to-host copy ordered on the engine's stream (not the legacy one)
void Engine::d2h(void *dst, const void *src, size_t bytes)
{
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
}

int64_t Engine::neighbor_count()
{
    if (last_total_ < 0) {
        std::vector<int> cnt(last_nat_);
        if (last_nat_) d2h(cnt.data(), count_.p, 4 * (size_t)last_nat_);
        last_total_ = 0;
        for (int c : cnt) last_total_ += c;
    }
    return last_total_;
}

void Engine::get_neighbors(int64_t cap, int32_t *ent)
{
    const int64_t total = neighbor_count();
    if (cap < total) throw Error(1, "neighbour buffer too small");
    if (total == 0) return;
    const size_t slots = (size_t)last_nat_ * last_cap_;
    std::vector<int> cnt(last_nat_), j(slots), sh(slots);
    d2h(cnt.data(), count_.p, 4 * (size_t)last_nat_);
    d2h(j.data(), ent_j_.p, 4 * slots);
    d2h(sh.data(), ent_shift_.p, 4 * slots);
    int64_t q = 0;
    for (int i = 0; i < last_nat_; ++i)
        for (int n = 0; n < cnt[i]; ++n, ++q) {
            const size_t t = (size_t)i * last_cap_ + n;
            int x, y, z;
            unpack_shift(sh[t], x, y, z);
            ent[5 * q + 0] = i;
            ent[5 * q + 1] = j[t];
            ent[5 * q + 2] = x;
            ent[5 * q + 3] = y;
            ent[5 * q + 4] = z;
        }
}

void Engine::get_symfunctions(int ld, double *G, double *dEdG, double *Ei)
{
    const int nat = last_nat_, ns = dm_.nsfmax;
    if (nat == 0) return;
    if ((G || dEdG) && ld < ns) throw Error(1, "ld smaller than the largest SF count");
    const bool fp64 = precision_ == 0;
    std::vector<double> tmp((size_t)nat * ns);
    auto fetch = [&](const DevMem &src, double *dst) {
        if (fp64) {
            d2h(tmp.data(), src.p, 8 * tmp.size());
        } else {
            std::vector<float> t(tmp.size());
            d2h(t.data(), src.p, 4 * t.size());
            for (size_t k = 0; k < t.size(); ++k) tmp[k] = t[k];
        }
        for (int i = 0; i < nat; ++i)
            for (int s = 0; s < ld; ++s) dst[(size_t)i * ld + s] = s < ns ? tmp[(size_t)i * ns + s] : 0.0;
    };
    if (G) fetch(G_, G);
    if (dEdG) fetch(dEdG_, dEdG);
    if (Ei) {   // committee mean of the atomic energies
        const int nm = (int)dmm_.size();
        std::vector<double> all((size_t)nat * nm);
        d2h(all.data(), Ei_.p, 8 * all.size());
        for (int i = 0; i < nat; ++i) {
            double a = 0;
            for (int k = 0; k < nm; ++k) a += all[(size_t)k * nat + i];
            Ei[i] = nm == 1 ? all[i] : a / nm;
        }
    }
}

void Engine::timings(double *ms6) const
{
    for (int k = 0; k < 6; ++k) ms6[k] = ms_[k];
}

}  // namespace hdnnp
