// C ABI (include/hdnnp_capi.h): thin extern "C" layer over hdnnp::Engine.
// Exceptions never cross the boundary; they become status codes plus a
// message retrievable with hdnnp_last_error().
#include "../../include/hdnnp_capi.h"

#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "engine.hpp"
#include "md.hpp"
#include "runner.hpp"
#include "server.hpp"

struct hdnnp_ctx {
    hdnnp::Engine *eng = nullptr;
    std::string err;
};

struct hdnnp_server {
    hdnnp::Server *sv = nullptr;
};

struct hdnnp_md {
    hdnnp_ctx *ctx = nullptr;     // owns the engine the systems are stepped with
    hdnnp::MdSystems *sys = nullptr;
};

static thread_local std::string g_create_error;
static thread_local std::string g_server_error;

// server calls: the message goes to the calling thread's server error
template <typename F>
static int sguarded(F &&f)
{
    try {
        f();
        return HDNNP_OK;
    } catch (const hdnnp::Error &e) {
        g_server_error = e.what();
        return e.code;
    } catch (const std::bad_alloc &) {
        g_server_error = "host out of memory";
        return HDNNP_ERR_CAPACITY;
    } catch (const std::exception &e) {
        g_server_error = e.what();
        return HDNNP_ERR_DEVICE;
    }
}

template <typename F>
static int guarded(hdnnp_ctx *ctx, F &&f)
{
    try {
        f();
        if (ctx) ctx->err.clear();
        return HDNNP_OK;
    } catch (const hdnnp::Error &e) {
        (ctx ? ctx->err : g_create_error) = e.what();
        return e.code;
    } catch (const std::bad_alloc &) {
        (ctx ? ctx->err : g_create_error) = "host out of memory";
        return HDNNP_ERR_CAPACITY;
    } catch (const std::exception &e) {
        (ctx ? ctx->err : g_create_error) = e.what();
        return HDNNP_ERR_DEVICE;
    }
}

// A context's device copy of the potential is set up when it is created if a
// GPU is visible (else on its first evaluation), so that an evaluation that
// may not allocate (hdnnp_evaluate_batch_device_noalloc) finds it there.
static void set_up_device(hdnnp::Engine &eng)
{
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) == hipSuccess && ndev > 0) eng.ensure_device();
    (void)hipGetLastError();
}

extern "C" {

int hdnnp_create(const char *param_dir, int32_t precision, hdnnp_ctx **out)
{
    if (!out || !param_dir) {
        g_create_error = "hdnnp_create: null argument";
        return HDNNP_ERR_ARG;
    }
    *out = nullptr;
    return guarded(nullptr, [&] {
        std::vector<hdnnp::HostModel> ms;
        ms.push_back(hdnnp::read_runner(param_dir));
        auto *c = new hdnnp_ctx;
        try {
            c->eng = new hdnnp::Engine(std::move(ms), precision);
            set_up_device(*c->eng);
        } catch (...) {
            delete c->eng;
            delete c;
            throw;
        }
        *out = c;
    });
}

int hdnnp_create_committee(int32_t n, const char *const *param_dirs, int32_t precision, hdnnp_ctx **out)
{
    if (!out || !param_dirs || n < 1) {
        g_create_error = "hdnnp_create_committee: null argument or n < 1";
        return HDNNP_ERR_ARG;
    }
    *out = nullptr;
    return guarded(nullptr, [&] {
        std::vector<hdnnp::HostModel> ms;
        for (int k = 0; k < n; ++k) {
            if (!param_dirs[k]) throw hdnnp::Error(HDNNP_ERR_ARG, "hdnnp_create_committee: null directory");
            ms.push_back(hdnnp::read_runner(param_dirs[k]));
        }
        auto *c = new hdnnp_ctx;
        try {
            c->eng = new hdnnp::Engine(std::move(ms), precision);
            set_up_device(*c->eng);
        } catch (...) {
            delete c->eng;
            delete c;
            throw;
        }
        *out = c;
    });
}

int hdnnp_md_create(hdnnp_ctx *ctx, int32_t nsys, const int32_t *nat, const int32_t *Z, const double *pos,
                    const double *mom, const double *mass, const int32_t *periodic, const double *lat, double dt,
                    double skin, hdnnp_md **out)
{
    if (!ctx || !out) return HDNNP_ERR_ARG;
    *out = nullptr;
    return guarded(ctx, [&] {
        auto *m = new hdnnp_md;
        m->ctx = ctx;
        try {
            m->sys = new hdnnp::MdSystems(*ctx->eng, nsys, nat, Z, pos, mom, mass, periodic, lat, dt, skin);
        } catch (...) {
            delete m;
            throw;
        }
        *out = m;
    });
}

int hdnnp_md_step(hdnnp_md *md, int32_t nsteps, double *epot, double *ekin)
{
    if (!md) return HDNNP_ERR_ARG;
    return guarded(md->ctx, [&] { md->sys->step(nsteps, epot, ekin); });
}

int hdnnp_md_get(hdnnp_md *md, double *pos, double *mom, double *force)
{
    if (!md) return HDNNP_ERR_ARG;
    return guarded(md->ctx, [&] { md->sys->get(pos, mom, force); });
}

int hdnnp_md_stats(const hdnnp_md *md, int64_t *steps, int64_t *list_builds)
{
    if (!md) return HDNNP_ERR_ARG;
    if (steps) *steps = md->sys->steps();
    if (list_builds) *list_builds = md->sys->builds();
    return HDNNP_OK;
}

int hdnnp_md_counters(const hdnnp_md *md, int64_t *out)
{
    if (!md || !out) return HDNNP_ERR_ARG;
    out[0] = md->sys->steps();
    out[1] = md->sys->builds();
    out[2] = md->sys->rollbacks();
    out[3] = md->sys->graph_captures();
    out[4] = md->sys->graph_replays();
    return HDNNP_OK;
}

void hdnnp_md_destroy(hdnnp_md *md)
{
    if (!md) return;
    delete md->sys;
    delete md;
}

int hdnnp_committee_size(const hdnnp_ctx *ctx, int32_t *n)
{
    if (!ctx || !n) return HDNNP_ERR_ARG;
    *n = ctx->eng->committee_size();
    return HDNNP_OK;
}

int hdnnp_last_committee(hdnnp_ctx *ctx, int32_t nstruct, double *std_dev, double *member_energies)
{
    if (!ctx) return HDNNP_ERR_ARG;
    return guarded(ctx, [&] {
        const std::vector<double> &sd = ctx->eng->energy_std();
        const std::vector<double> &me = ctx->eng->member_energies();
        if (nstruct != (int32_t)sd.size())
            throw hdnnp::Error(HDNNP_ERR_ARG, "nstruct differs from the last evaluation's (" +
                                                  std::to_string(sd.size()) + ")");
        if (std_dev) std::copy(sd.begin(), sd.end(), std_dev);
        if (member_energies) std::copy(me.begin(), me.end(), member_energies);
    });
}

int hdnnp_evaluate(hdnnp_ctx *ctx, int32_t nat, const int32_t *Z, const double *pos, int32_t periodic,
                   const double *lat, double *energy, double *force, double *dEdLat)
{
    if (!ctx || !energy) return HDNNP_ERR_ARG;
    return guarded(ctx, [&] { ctx->eng->evaluate_host(nat, Z, pos, periodic, lat, energy, force, dEdLat); });
}

int hdnnp_evaluate_batch(hdnnp_ctx *ctx, int32_t nstruct, const int32_t *nat, const int32_t *Z, const double *pos,
                         const int32_t *periodic, const double *lat, double *energy, double *force, double *dEdLat)
{
    if (!ctx) return HDNNP_ERR_ARG;
    return guarded(ctx, [&] {
        ctx->eng->evaluate_batch_host(nstruct, nat, Z, pos, periodic, lat, energy, force, dEdLat);
    });
}

int hdnnp_evaluate_batch_device(hdnnp_ctx *ctx, int32_t nstruct, const int32_t *nat, const int32_t *Z_dev,
                                const double *pos_dev, const int32_t *periodic, const double *lat, double *energy,
                                double *force_dev, double *dEdLat, void *stream)
{
    if (!ctx) return HDNNP_ERR_ARG;
    return guarded(ctx, [&] {
        ctx->eng->evaluate_batch_device(nstruct, nat, Z_dev, pos_dev, periodic, lat, energy, force_dev, dEdLat,
                                        (hipStream_t)stream);
    });
}

int hdnnp_evaluate_batch_device_noalloc(hdnnp_ctx *ctx, int32_t nstruct, const int32_t *nat, const int32_t *Z_dev,
                                        const double *pos_dev, const int32_t *periodic, const double *lat,
                                        double *energy, double *force_dev, double *dEdLat, void *stream)
{
    if (!ctx) return HDNNP_ERR_ARG;
    return guarded(ctx, [&] {
        ctx->eng->evaluate_batch_device_noalloc(nstruct, nat, Z_dev, pos_dev, periodic, lat, energy, force_dev,
                                                dEdLat, (hipStream_t)stream);
    });
}

int hdnnp_evaluate_device(hdnnp_ctx *ctx, int32_t nat, const int32_t *Z_dev, const double *pos_dev,
                          int32_t periodic, const double *lat, double *energy, double *force_dev,
                          double *dEdLat, void *stream)
{
    if (!ctx || !energy) return HDNNP_ERR_ARG;
    return guarded(ctx, [&] {
        ctx->eng->evaluate_device(nat, Z_dev, pos_dev, periodic, lat, energy, force_dev, dEdLat,
                                  (hipStream_t)stream);
    });
}

// the stage's lists are the engine's current ones, else they are built again
// from the structure given
static void stage_ensure(hdnnp_ctx *ctx, int64_t *stage, int32_t nat, const int32_t *Z, const double *pos,
                         int32_t periodic, const double *lat)
{
    if (!ctx->eng->stage_current(*stage)) *stage = ctx->eng->stage_build(nat, Z, pos, periodic, lat);
}

int hdnnp_stage_neighbor_lists(hdnnp_ctx *ctx, int32_t nat, const int32_t *Z, const double *pos, int32_t periodic,
                               const double *lat, int64_t *stage, int32_t *nneis)
{
    if (!ctx || !stage || (nat > 0 && !nneis)) return HDNNP_ERR_ARG;
    return guarded(ctx, [&] {
        *stage = ctx->eng->stage_build(nat, Z, pos, periodic, lat);
        ctx->eng->stage_neighbors(nneis, nullptr, nullptr);
    });
}

int hdnnp_stage_get_neighbor_lists(hdnnp_ctx *ctx, int64_t stage, int32_t *ineis, double *neis)
{
    if (!ctx) return HDNNP_ERR_ARG;
    return guarded(ctx, [&] {
        if (!ctx->eng->stage_current(stage))
            throw hdnnp::Error(HDNNP_ERR_ARG, "the stage is not the context's current one (it built lists since)");
        std::vector<int32_t> nn(std::max(ctx->eng->prepared_nat(), 1));
        ctx->eng->stage_neighbors(nn.data(), ineis, neis);
    });
}

int hdnnp_stage_symfunctions(hdnnp_ctx *ctx, int64_t *stage, int32_t nat, const int32_t *Z, const double *pos,
                             int32_t periodic, const double *lat, int32_t together, int32_t ld, double *G, double *dsf)
{
    if (!ctx || !stage || (nat > 0 && !G)) return HDNNP_ERR_ARG;
    return guarded(ctx, [&] {
        stage_ensure(ctx, stage, nat, Z, pos, periodic, lat);
        ctx->eng->stage_symfunctions(together, ld, G, dsf);
    });
}

int hdnnp_stage_nns(hdnnp_ctx *net, hdnnp_ctx *sfctx, int64_t *stage, int32_t nat, const int32_t *Z, const double *pos,
                    int32_t periodic, const double *lat, int32_t ld, const double *G, double *energy, double *force,
                    double *dEdLat)
{
    if (!net || !stage || !energy || (nat > 0 && (!G || !force))) return HDNNP_ERR_ARG;
    hdnnp_ctx *sc = sfctx ? sfctx : net;
    const int rc = guarded(sc, [&] {
        stage_ensure(sc, stage, nat, Z, pos, periodic, lat);
        sc->eng->stage_nns(*net->eng, ld, G, energy, force, dEdLat);
    });
    if (rc && sc != net) net->err = sc->err;
    return rc;
}

int hdnnp_neighbor_count(hdnnp_ctx *ctx, int64_t *count)
{
    if (!ctx || !count) return HDNNP_ERR_ARG;
    *count = ctx->eng->neighbor_count();
    return HDNNP_OK;
}

int hdnnp_get_neighbors(hdnnp_ctx *ctx, int64_t cap, int32_t *ent)
{
    if (!ctx || (!ent && cap > 0)) return HDNNP_ERR_ARG;
    return guarded(ctx, [&] { ctx->eng->get_neighbors(cap, ent); });
}

int hdnnp_get_symfunctions(hdnnp_ctx *ctx, int32_t ld, double *G, double *dEdG, double *Ei)
{
    if (!ctx) return HDNNP_ERR_ARG;
    return guarded(ctx, [&] { ctx->eng->get_symfunctions(ld, G, dEdG, Ei); });
}

int hdnnp_get_model_info(hdnnp_ctx *ctx, hdnnp_model_info *info)
{
    if (!ctx || !info) return HDNNP_ERR_ARG;
    const hdnnp::HostModel &m = ctx->eng->model();
    std::memset(info, 0, sizeof *info);
    info->nelements = (int32_t)m.el.size();
    for (size_t e = 0; e < m.el.size() && e < 16; ++e) {
        info->Z[e] = m.el[e].Z;
        info->nsf[e] = (int32_t)m.el[e].sfs.size();
        info->atom_energy[e] = m.el[e].atom_energy;
    }
    if (!m.el.empty()) {
        info->nlayers = (int32_t)m.el[0].layers.size();
        for (size_t l = 0; l < m.el[0].layers.size() && l < 8; ++l) {
            info->nodes[l] = m.el[0].layers[l].nout;
            info->act[l] = m.el[0].layers[l].act;
        }
    }
    info->scalesf = m.scalesf;
    info->centersf = m.centersf;
    info->maxrc = m.maxrc;
    info->emin = m.emin;
    info->emax = m.emax;
    info->cutoff_type = m.cutoff_type;
    info->scale_range = m.scale_range;
    info->scale_min = m.scale_min;
    info->scale_max = m.scale_max;
    return HDNNP_OK;
}

int hdnnp_get_sf_table(hdnnp_ctx *ctx, int32_t e, int32_t cap_rows, double *rows)
{
    if (!ctx || !rows) return HDNNP_ERR_ARG;
    const hdnnp::HostModel &m = ctx->eng->model();
    if (e < 0 || e >= (int)m.el.size()) {
        ctx->err = "element rank out of range";
        return HDNNP_ERR_ARG;
    }
    const auto &sfs = m.el[e].sfs;
    if (cap_rows < (int)sfs.size()) {
        ctx->err = "row buffer too small";
        return HDNNP_ERR_ARG;
    }
    for (size_t s = 0; s < sfs.size(); ++s) {
        const auto &f = sfs[s];
        double *r = rows + 11 * s;
        r[0] = f.type;
        r[1] = f.e1;
        r[2] = f.type != 2 ? f.e2 : 0;
        r[3] = f.rcut;
        r[4] = f.eta;
        r[5] = f.type == 2 ? f.rs : 0.0;
        r[6] = f.type != 2 ? f.lambda : 0.0;
        r[7] = f.type != 2 ? f.zeta : 0.0;
        r[8] = f.sfmin;
        r[9] = f.sfmax;
        r[10] = f.sfmean;
    }
    return HDNNP_OK;
}

int hdnnp_get_weights(hdnnp_ctx *ctx, int32_t e, int64_t cap, double *w, int64_t *count)
{
    if (!ctx || !count) return HDNNP_ERR_ARG;
    const hdnnp::HostModel &m = ctx->eng->model();
    if (e < 0 || e >= (int)m.el.size()) {
        ctx->err = "element rank out of range";
        return HDNNP_ERR_ARG;
    }
    int64_t n = 0;
    for (auto &L : m.el[e].layers) n += (int64_t)L.W.size() + (int64_t)L.b.size();
    *count = n;
    if (!w) return HDNNP_OK;
    if (cap < n) {
        ctx->err = "weight buffer too small";
        return HDNNP_ERR_ARG;
    }
    int64_t k = 0;
    for (auto &L : m.el[e].layers) {
        for (double v : L.W) w[k++] = v;
        for (double v : L.b) w[k++] = v;
    }
    return HDNNP_OK;
}

int hdnnp_dEdLat_to_stress(const double *lat, const double *dEdLat, double *stress)
{
    if (!lat || !dEdLat || !stress) return HDNNP_ERR_ARG;
    // stress = -1/det3D(lat) * matmul(dEdLat, transpose(lat)); Fortran A(r,c) at [r + 3c]
    const double det = hdnnp::det3(lat);
    double out[9];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            double acc = 0.0;
            for (int k = 0; k < 3; ++k) acc += dEdLat[r + 3 * k] * lat[c + 3 * k];
            out[r + 3 * c] = -1.0 / det * acc;
        }
    std::memcpy(stress, out, sizeof out);
    return HDNNP_OK;
}

const char *hdnnp_last_error(const hdnnp_ctx *ctx)
{
    return ctx ? ctx->err.c_str() : g_create_error.c_str();
}

int hdnnp_last_timings(hdnnp_ctx *ctx, double *ms6)
{
    if (!ctx || !ms6) return HDNNP_ERR_ARG;
    ctx->eng->timings(ms6);
    return HDNNP_OK;
}

int hdnnp_set_stage_timings(hdnnp_ctx *ctx, int32_t on)
{
    if (!ctx || (on != 0 && on != 1)) return HDNNP_ERR_ARG;
    return guarded(ctx, [&] { ctx->eng->set_stage_timings(on != 0); });
}

int hdnnp_set_graphs(hdnnp_ctx *ctx, int32_t on)
{
    if (!ctx || (on != 0 && on != 1)) return HDNNP_ERR_ARG;
    return guarded(ctx, [&] { ctx->eng->set_eval_graphs(on != 0); });
}

int hdnnp_graph_counts(const hdnnp_ctx *ctx, int32_t *on, int64_t *captures, int64_t *replays)
{
    if (!ctx) return HDNNP_ERR_ARG;
    if (on) *on = ctx->eng->eval_graphs() ? 1 : 0;
    ctx->eng->eval_graph_counts(captures, replays);
    return HDNNP_OK;
}

void hdnnp_destroy(hdnnp_ctx *ctx)
{
    if (!ctx) return;
    delete ctx->eng;
    delete ctx;
}

int hdnnp_set_tile(hdnnp_ctx *ctx, int32_t atoms_per_wave)
{
    if (!ctx) return HDNNP_ERR_ARG;
    return guarded(ctx, [&] { ctx->eng->set_tile(atoms_per_wave); });
}

int hdnnp_get_tile(const hdnnp_ctx *ctx, int32_t *atoms_per_wave)
{
    if (!ctx || !atoms_per_wave) return HDNNP_ERR_ARG;
    *atoms_per_wave = ctx->eng->tile();
    return HDNNP_OK;
}

int hdnnp_autotune(hdnnp_ctx *ctx, int32_t nstruct, const int32_t *nat, const int32_t *Z, const double *pos,
                   const int32_t *periodic, const double *lat, int32_t *atoms_per_wave, double *ms)
{
    if (!ctx) return HDNNP_ERR_ARG;
    return guarded(ctx, [&] {
        const int t = ctx->eng->autotune(nstruct, nat, Z, pos, periodic, lat, ms);
        if (atoms_per_wave) *atoms_per_wave = t;
    });
}

int hdnnp_server_create(int32_t n, const char *const *param_dirs, int32_t precision, int32_t lanes,
                        int64_t hbm_budget_bytes, int64_t max_batch_atoms, hdnnp_server **out)
{
    if (!out || !param_dirs || n < 1 || hbm_budget_bytes < 0) {
        g_server_error = "hdnnp_server_create: null argument, n < 1 or negative budget";
        return HDNNP_ERR_ARG;
    }
    *out = nullptr;
    return sguarded([&] {
        std::vector<hdnnp::HostModel> ms;
        for (int k = 0; k < n; ++k) {
            if (!param_dirs[k]) throw hdnnp::Error(HDNNP_ERR_ARG, "hdnnp_server_create: null directory");
            ms.push_back(hdnnp::read_runner(param_dirs[k]));
        }
        auto *h = new hdnnp_server;
        try {
            h->sv = new hdnnp::Server(std::move(ms), precision, lanes, (size_t)hbm_budget_bytes, (long)max_batch_atoms);
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

int hdnnp_server_submit(hdnnp_server *sv, int32_t nat, const int32_t *Z, const double *pos, int32_t periodic,
                        const double *lat, int64_t *ticket)
{
    if (!sv || !ticket) return HDNNP_ERR_ARG;
    return sguarded([&] { *ticket = sv->sv->submit(nat, Z, pos, periodic, lat); });
}

int hdnnp_server_wait(hdnnp_server *sv, int64_t ticket, double timeout_s, double *energy, double *force,
                      double *dEdLat, double *energy_std, double *latency_ms)
{
    if (!sv) return HDNNP_ERR_ARG;
    bool done = false;
    const int rc = sguarded([&] { done = sv->sv->wait(ticket, timeout_s, energy, force, dEdLat, energy_std, latency_ms); });
    if (rc == HDNNP_OK && !done) {
        g_server_error = "timed out";
        return HDNNP_ERR_TIMEOUT;
    }
    return rc;
}

int hdnnp_server_cancel(hdnnp_server *sv, int64_t ticket, int32_t *cancelled)
{
    if (!sv || !cancelled) return HDNNP_ERR_ARG;
    return sguarded([&] { *cancelled = sv->sv->cancel(ticket) ? 1 : 0; });
}

int hdnnp_server_md_open(hdnnp_server *sv, int32_t nat, const int32_t *Z, const double *pos, const double *mom,
                         const double *mass, int32_t periodic, const double *lat, double dt, double skin,
                         int64_t *session)
{
    if (!sv || !session) return HDNNP_ERR_ARG;
    return sguarded([&] { *session = sv->sv->md_open(nat, Z, pos, mom, mass, periodic, lat, dt, skin); });
}

int hdnnp_server_md_step(hdnnp_server *sv, int64_t session, int32_t nsteps, int64_t *ticket)
{
    if (!sv || !ticket) return HDNNP_ERR_ARG;
    return sguarded([&] { *ticket = sv->sv->md_step(session, nsteps); });
}

int hdnnp_server_md_wait(hdnnp_server *sv, int64_t ticket, double timeout_s, double *epot, double *ekin,
                         double *latency_ms)
{
    if (!sv) return HDNNP_ERR_ARG;
    bool done = false;
    const int rc = sguarded([&] { done = sv->sv->md_wait(ticket, timeout_s, epot, ekin, latency_ms); });
    if (rc == HDNNP_OK && !done) {
        g_server_error = "timed out";
        return HDNNP_ERR_TIMEOUT;
    }
    return rc;
}

int hdnnp_server_md_get(hdnnp_server *sv, int64_t session, double *pos, double *mom, double *force)
{
    if (!sv) return HDNNP_ERR_ARG;
    return sguarded([&] { sv->sv->md_get(session, pos, mom, force); });
}

int hdnnp_server_md_close(hdnnp_server *sv, int64_t session)
{
    if (!sv) return HDNNP_ERR_ARG;
    return sguarded([&] { sv->sv->md_close(session); });
}

int hdnnp_server_stats(hdnnp_server *sv, int32_t reset, double *out)
{
    if (!sv || !out) return HDNNP_ERR_ARG;
    return sguarded([&] { sv->sv->stats(out, reset != 0); });
}

const char *hdnnp_server_last_error(void) { return g_server_error.c_str(); }

void hdnnp_server_destroy(hdnnp_server *sv)
{
    if (!sv) return;
    delete sv->sv;
    delete sv;
}

}  // extern "C"
