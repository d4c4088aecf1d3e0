/* hdnnp_capi.h -- C ABI of the MI355X HDNNP engine (libhdnnp_amd.so).
 *
 * The Fortran module `hdnnps` of this package binds these entry points with
 * iso_c_binding; any C/C++/Python caller can use them directly.
 *
 * Conventions (those of the reference's atStruct, atomicStructure.f90:17-28):
 *   pos[nat][3]   bohr, x1 y1 z1 x2 ...           (Fortran ats(3,nat))
 *   lat[3][3]     row k = lattice vector k          (Fortran lat(:,k))
 *   Z[nat]        atomic numbers, int32             (Fortran el(nat))
 *   force[nat][3] Ha/bohr, same layout as pos
 *   dEdLat[3][3]  dEdLat[k][a] = dE/d lat(a,k) ... i.e. the same 9 doubles
 *                 as Fortran dEdLat(3,3) in memory order
 *   energies in Hartree.
 * Every entry point returns 0 on success and a non-zero status otherwise;
 * hdnnp_last_error() then holds a message. Calls on one context must not
 * overlap; separate contexts may be used concurrently.
 */
#ifndef HDNNP_CAPI_H
#define HDNNP_CAPI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hdnnp_ctx hdnnp_ctx;

/* arithmetic of an evaluation */
enum {
    HDNNP_FP64 = 0,        /* everything in FP64                                 */
    HDNNP_MIXED = 1,       /* wrap/list/energy/virial FP64, symmetry functions   */
                           /* FP32, MLP FP32                                     */
    HDNNP_MIXED_BF16 = 2   /* as HDNNP_MIXED, with BF16 operands (FP32           */
                           /* accumulation) in the MLP's forward layers; its     */
                           /* backward pass stays FP32                           */
};

/* status codes */
enum {
    HDNNP_OK = 0,
    HDNNP_ERR_ARG = 1,      /* bad argument                                */
    HDNNP_ERR_IO = 2,       /* parameter file missing or malformed         */
    HDNNP_ERR_MODEL = 3,    /* model not usable (unsorted/mixed-rc SFs...) */
    HDNNP_ERR_ELEMENT = 4,  /* structure holds an element the model lacks  */
    HDNNP_ERR_DEVICE = 5,   /* HIP runtime error                           */
    HDNNP_ERR_CAPACITY = 6, /* a size limit of this build was exceeded     */
    HDNNP_ERR_TIMEOUT = 7,  /* a server wait ran out of time               */
    HDNNP_ERR_CANCELLED = 8 /* the server request was cancelled            */
};

/* Replaces parseRuNNerInput(handle, path) (src/potentials/hdnnps.f90:34-233):
 * reads path/input.nn, path/scaling.data, path/weights.ZZZ.data (param_dir
 * is the directory; a trailing '/' is optional). Parsing needs no GPU. When
 * a GPU is visible, the device copy of the potential is set up here; else on
 * the first evaluation. */
int hdnnp_create(const char *param_dir, int32_t precision, hdnnp_ctx **out);

/* Replaces initGlobalHdnnpsFromInputFile (src/potentials/globalHdnnps.f90:13-29,
 * which reads the directories from nnps.txt): a committee of n potentials.
 * The first potential's symmetry functions and scaling serve all members;
 * member k contributes its networks and atom energies (hdnnpCalcNNs with the
 * first handle's symmetry functions, globalHdnnps.f90:47-55). Evaluations of
 * the context (hdnnp_evaluate and its variants) then return the committee
 * mean of E, F and dEdLat (globalHdnnpEnergyAndForces, :31-70). n = 1 is the
 * same as hdnnp_create. */
int hdnnp_create_committee(int32_t n, const char *const *param_dirs, int32_t precision, hdnnp_ctx **out);
int hdnnp_committee_size(const hdnnp_ctx *ctx, int32_t *n);

/* Of the last evaluation (nstruct structures; 1 for the single-structure
 * calls): std_dev[nstruct] = sqrt(mean over members of (E_k - mean)^2)
 * (the `std` argument of globalHdnnpEnergyAndForces) and
 * member_energies[nstruct][n]. Either pointer may be NULL. */
int hdnnp_last_committee(hdnnp_ctx *ctx, int32_t nstruct, double *std_dev, double *member_energies);

/* Replaces hdnnpEnergyAndForces(handle, ats, energy, force, dEdLat)
 * (src/potentials/hdnnps.f90:333-371). Host arrays in, host arrays out.
 * dEdLat may be NULL; for nonperiodic input it is set to zero (the reference
 * leaves it undefined). */
int hdnnp_evaluate(hdnnp_ctx *ctx, int32_t nat, const int32_t *Z, const double *pos,
                   int32_t periodic, const double *lat, double *energy, double *force,
                   double *dEdLat);

/* Many independent structures in one evaluation (the reference evaluates one
 * structure per hdnnpEnergyAndForces call; this is that call for nstruct
 * structures at once). Structure k has nat[k] atoms; Z, pos and force hold the
 * atoms of all structures one after the other (pos/force [sum nat][3]);
 * periodic[k]; lat[9k..9k+8] is read for periodic structures only (lat may be
 * NULL when none is periodic); energy[nstruct]; dEdLat[9k..] per structure
 * (may be NULL). Each result equals the result of hdnnp_evaluate on that
 * structure alone, within rounding of the summation order of E. */
int hdnnp_evaluate_batch(hdnnp_ctx *ctx, int32_t nstruct, const int32_t *nat, const int32_t *Z,
                         const double *pos, const int32_t *periodic, const double *lat, double *energy,
                         double *force, double *dEdLat);

/* hdnnp_evaluate_batch with Z_dev, pos_dev and force_dev resident in device
 * memory (nat, periodic, lat, energy and dEdLat are host arrays). `stream` as
 * for hdnnp_evaluate_device. No counterpart in the reference. */
int hdnnp_evaluate_batch_device(hdnnp_ctx *ctx, int32_t nstruct, const int32_t *nat, const int32_t *Z_dev,
                                const double *pos_dev, const int32_t *periodic, const double *lat,
                                double *energy, double *force_dev, double *dEdLat, void *stream);

/* hdnnp_evaluate_batch_device for callers that cannot allocate: code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc.
 * No counterpart in the reference. Same arguments and results (bit for bit)
 * as hdnnp_evaluate_batch_device, but the call makes no device or pinned
 * allocation and frees nothing: its work buffers are carved from a
 * file-scope device array of HDNNP_SCRATCH_BYTES (128 MiB) that every GPU the
 * library runs on holds until the process ends, host staging uses pageable
 * memory, and the flags are cleared by a fill kernel. The context's device
 * copy of the potential must exist already: hdnnp_create sets it up when a
 * GPU is visible (otherwise the call fails with HDNNP_ERR_DEVICE). Inputs
 * needing more scratch than the array holds are refused with
 * HDNNP_ERR_CAPACITY before anything is launched (a cell of the 80-SF
 * perovskite potential fits up to about 10,000 atoms in mixed precision,
 * 8,400 in FP64). All calls on
 * one GPU share the array, which keeps the last call's contents; the library
 * serialises such calls (any contexts, any threads), so overlapping calls
 * wait for each other instead of racing. The context's own work buffers are
 * left as they were; stage outputs (hdnnp_get_neighbors, ...) of the call
 * are not kept. */
#define HDNNP_SCRATCH_BYTES (128u << 20)
int hdnnp_evaluate_batch_device_noalloc(hdnnp_ctx *ctx, int32_t nstruct, const int32_t *nat, const int32_t *Z_dev,
                                        const double *pos_dev, const int32_t *periodic, const double *lat,
                                        double *energy, double *force_dev, double *dEdLat, void *stream);

/* Same evaluation with Z, pos and force resident in device memory (HBM);
 * lat, energy and dEdLat are host pointers. `stream` is a hipStream_t or
 * NULL for the context's own stream; the call returns after the results are
 * complete. No counterpart in the reference (which has no device). */
int hdnnp_evaluate_device(hdnnp_ctx *ctx, int32_t nat, const int32_t *Z_dev,
                          const double *pos_dev, int32_t periodic, const double *lat,
                          double *energy, double *force_dev, double *dEdLat, void *stream);

/* The reference's stage API (hdnnps.f90:373-562): hdnnpCalcNeighborLists,
 * hdnnpCalcSymfunctions and hdnnpCalcNNs, which globalHdnnps.f90:47-57 calls
 * to share one potential's lists and symmetry functions among several
 * potentials' networks. A context keeps one stage (the lists of one
 * structure) on the device; `stage` names it. Every call but
 * hdnnp_stage_get_neighbor_lists takes the structure again (as the
 * reference's routines take ats): when the context has built other lists
 * since, the stage's lists are built again from it and *stage is renewed.
 * Structure arguments as for hdnnp_evaluate.
 *
 * hdnnpCalcNeighborLists (hdnnps.f90:373-388): builds the lists;
 * nneis[nat] = entries per atom with the atom itself in slot 1
 * (neighborLists.f90:118-127). */
int hdnnp_stage_neighbor_lists(hdnnp_ctx *ctx, int32_t nat, const int32_t *Z, const double *pos, int32_t periodic,
                               const double *lat, int64_t *stage, int32_t *nneis);
/* The lists of the context's current stage in the reference's layout and
 * order (type neighborList, neighborLists.f90:10-16), atom after atom:
 * ineis[sum nneis] 1-based atom numbers, neis[sum nneis][3] the positions of
 * the images (slot 1 the atom as the lists were built from it: wrapped into
 * the cell when periodic). Either pointer may be NULL. */
int hdnnp_stage_get_neighbor_lists(hdnnp_ctx *ctx, int64_t stage, int32_t *ineis, double *neis);
/* hdnnpCalcSymfunctions (hdnnps.f90:390-485): scaled G[nat][ld] (ld >= the
 * largest SF count; columns an element lacks are 0) and, when dsf is not
 * NULL, the scaled dG/dR of the reference's symFunctionContainer
 * (symmetryFunctions.f90:24-33), atom after atom, function after function,
 * slot after slot: dsf[..][3] for the nneis slots of the atom. together != 0
 * takes G from the fused pass (calcAllSymmetryFunctions,
 * symmetryFunctions.f90:158-308), 0 from the per-function kernel
 * (symmetryFunction2 / symmetryFunction3, :41-152); dsf always comes from the
 * per-function kernel. */
int hdnnp_stage_symfunctions(hdnnp_ctx *ctx, int64_t *stage, int32_t nat, const int32_t *Z, const double *pos,
                             int32_t periodic, const double *lat, int32_t together, int32_t ld, double *G, double *dsf);
/* hdnnpCalcNNs (hdnnps.f90:487-562): the networks and atom energies of `net`
 * (one potential) on the symmetry functions G[nat][ld] given and the lists
 * of sfctx's stage (NULL: net's): energy, force[nat][3], dEdLat (may be
 * NULL). The forces and dEdLat contract dE/dG with the gradients of sfctx's
 * symmetry functions, recomputed on the device from its lists. Both contexts
 * need the same precision, elements and SF counts. */
int hdnnp_stage_nns(hdnnp_ctx *net, hdnnp_ctx *sfctx, int64_t *stage, int32_t nat, const int32_t *Z, const double *pos,
                    int32_t periodic, const double *lat, int32_t ld, const double *G, double *energy, double *force,
                    double *dEdLat);

/* Stage outputs of the last evaluation, for the reference's stage API:
 * hdnnpCalcNeighborLists (hdnnps.f90:373-388) -> neighbour entries
 * (i, j, Lx, Ly, Lz), 0-based atoms, L the integer image of j, self slot
 * excluded, within the potential's largest cutoff. */
int hdnnp_neighbor_count(hdnnp_ctx *ctx, int64_t *count);
int hdnnp_get_neighbors(hdnnp_ctx *ctx, int64_t cap, int32_t *ent);

/* hdnnpCalcSymfunctions (hdnnps.f90:390-485) and the dE/dG of hdnnpCalcNNs
 * (hdnnps.f90:524-531) of the last evaluation: G[nat][ld] centred/scaled
 * symmetry functions, dEdG[nat][ld] = dE_i/dG_s (w.r.t. the scaled G), Ei[nat]
 * atomic energies (NN output + atom_energy). Any pointer may be NULL. */
int hdnnp_get_symfunctions(hdnnp_ctx *ctx, int32_t ld, double *G, double *dEdG, double *Ei);

typedef struct {
    int32_t nelements;
    int32_t Z[16];
    int32_t nsf[16];
    int32_t nlayers;        /* weight layers incl. output                */
    int32_t nodes[8];       /* nOut per layer                           */
    char act[8];            /* 'l' 't' 'p' 'r'                           */
    int32_t scalesf, centersf;
    double maxrc;
    double atom_energy[16];
    double emin, emax;      /* scaling.data energy range (parsed, unused) */
    /* RuNNer keywords the reference does not read (it always uses
     * cutoff_type 2 and centre/scale as below): cutoff_type 0..8 (default
     * 2); with scale_min_short / scale_max_short present (scale_range = 1)
     * a scaled SF is scale_min + (scale_max - scale_min) (G - c) / (max -
     * min), c = mean when centring, else min */
    int32_t cutoff_type, scale_range;
    double scale_min, scale_max;
} hdnnp_model_info;

int hdnnp_get_model_info(hdnnp_ctx *ctx, hdnnp_model_info *info);

/* Sorted symmetry-function table of element rank e (sortSymFunctions,
 * hdnnps.f90:235-267; readScaling :269-300): rows of 11 doubles
 * (type, Z1, Z2, rcut, eta, Rs, lambda, zeta, min, max, mean). */
int hdnnp_get_sf_table(hdnnp_ctx *ctx, int32_t e, int32_t cap_rows, double *rows);

/* Weights of element rank e in the reference's file order (readWeights,
 * hdnnps.f90:318-327): per layer W(i,j) input-major, then the biases. */
int hdnnp_get_weights(hdnnp_ctx *ctx, int32_t e, int64_t cap, double *w, int64_t *count);

/* Molecular dynamics on the device: the velocity-Verlet loop of the
 * reference's example (hdnnpExample.f90:58-72) for nsys systems at once, with
 * positions and momenta resident in device memory. nat[k] atoms per system;
 * Z, pos, mom [sum nat][3] (bohr, atomic momentum units) and mass[sum nat]
 * (electron masses) over all atoms; periodic[k] and lat[9k..] as for
 * hdnnp_evaluate_batch; dt in atomic time units. skin (bohr) > 0: neighbour
 * lists within rc + skin are reused while no atom has moved more than skin/2
 * since their build (a Verlet skin; the reference rebuilds at every step,
 * which skin = 0 does). The systems use the context's engine: do not
 * evaluate on the context while a step runs. Any other build on the context
 * (an evaluation, another MD object's step) replaces the lists, and the next
 * step builds them again: give MD objects that should reuse their lists a
 * context each. Forces of the initial positions are computed by
 * hdnnp_md_create. */
typedef struct hdnnp_md hdnnp_md;
int hdnnp_md_create(hdnnp_ctx *ctx, int32_t nsys, const int32_t *nat, const int32_t *Z, const double *pos,
                    const double *mom, const double *mass, const int32_t *periodic, const double *lat, double dt,
                    double skin, hdnnp_md **out);
/* nsteps steps; epot[nsys] and ekin[nsys] (= sum p^2/m / 2) of the last one (may be NULL) */
int hdnnp_md_step(hdnnp_md *md, int32_t nsteps, double *epot, double *ekin);
int hdnnp_md_get(hdnnp_md *md, double *pos, double *mom, double *force);
int hdnnp_md_stats(const hdnnp_md *md, int64_t *steps, int64_t *list_builds);
/* out[5]: steps, list builds, chunks cut back at a needed build, captures
 * of the reused-list step graph, replays of it. Steps on reused lists run in
 * chunks checked once at their end; all but a chunk's first step are
 * replays of one captured HIP graph of a step (HDNNP_MD_GRAPHS=0 in the
 * environment launches them directly). The trajectory is the same either
 * way, bit for bit. No counterpart in the reference. */
int hdnnp_md_counters(const hdnnp_md *md, int64_t *out);
void hdnnp_md_destroy(hdnnp_md *md);

/* Tile of the symmetry-function kernels, one per context (no counterpart
 * in the reference): atoms per wavefront, 1 (default), 2 or 4. Several atoms
 * per wavefront fill the lanes when atoms have few neighbours (molecules,
 * small clusters); one is fastest for condensed matter. Kernels serving more
 * than 128 staged neighbours, or whose LDS would not fit, use a smaller tile.
 * The tile sets the order in which partial G4 sums are added, so results of
 * different tiles agree to rounding. Batched results equal single ones bit
 * for bit when both run the same tile: always for tiles 1 and 2; tile 4
 * falls back to 2 in FP64 beyond 64 staged neighbours (LDS). */
int hdnnp_set_tile(hdnnp_ctx *ctx, int32_t atoms_per_wave);
int hdnnp_get_tile(const hdnnp_ctx *ctx, int32_t *atoms_per_wave);
/* Autotuner: evaluates a sample (as for hdnnp_evaluate_batch) with each tile,
 * times the two symmetry-function passes by HIP events (best of 3 after a
 * warm-up), and keeps the fastest as the context's tile. atoms_per_wave gets
 * the choice, ms[3] (may be NULL) the time of tiles 1, 2, 4 (0: not taken by
 * the sample's staging capacity). */
int hdnnp_autotune(hdnnp_ctx *ctx, int32_t nstruct, const int32_t *nat, const int32_t *Z, const double *pos,
                   const int32_t *periodic, const double *lat, int32_t *atoms_per_wave, double *ms);

/* Request server. The reference has no counterpart: its potential is called
 * by one serial caller at a time (hdnnps.f90:333, hdnnpExample.f90:58-72).
 * The server keeps the potential resident on the device for many callers:
 *  - Evaluations (hdnnp_server_submit / _wait) are queued and batched
 *    continuously. Each free lane takes every queued evaluation that fits,
 *    in arrival order, up to max_batch_atoms atoms (0: 262144), and
 *    evaluates them as one batch. Lanes are worker threads with their own
 *    stream and buffers, so their work overlaps on the device.
 *  - MD sessions (hdnnp_server_md_*) keep positions, momenta and neighbour
 *    lists (with a Verlet skin) resident between requests. Their requests
 *    run when no evaluation waits, or once they have waited 100 ms and no
 *    other lane runs an MD request.
 *  - Admission is by device memory. The buffers of the lanes and the state
 *    of the sessions stay within hbm_budget_bytes (0: 80% of the free
 *    device memory). Work waits until it fits; work larger than the whole
 *    budget is refused with HDNNP_ERR_CAPACITY.
 *  - Results are the same alone and in any mix: a structure's energy,
 *    forces and dEdLat equal those of hdnnp_evaluate bit for bit.
 * Units and layouts are those of hdnnp_evaluate and hdnnp_md_create. Every
 * ticket must be waited for once. Calls may come from any thread. Errors
 * of server calls are reported by hdnnp_server_last_error, per thread. */
typedef struct hdnnp_server hdnnp_server;
int hdnnp_server_create(int32_t n, const char *const *param_dirs, int32_t precision, int32_t lanes,
                        int64_t hbm_budget_bytes, int64_t max_batch_atoms, hdnnp_server **out);
/* queue one evaluation (the inputs are copied) */
int hdnnp_server_submit(hdnnp_server *sv, int32_t nat, const int32_t *Z, const double *pos, int32_t periodic,
                        const double *lat, int64_t *ticket);
/* Wait for an evaluation's results (timeout_s <= 0: no limit;
 * HDNNP_ERR_TIMEOUT leaves the ticket valid). Any output may be NULL;
 * energy_std is the committee's energy spread (0 for one potential);
 * latency_ms is the time from submission to completion. */
int hdnnp_server_wait(hdnnp_server *sv, int64_t ticket, double timeout_s, double *energy, double *force,
                      double *dEdLat, double *energy_std, double *latency_ms);
/* Cancel a queued evaluation or MD-step request (no counterpart in the
 * reference: the server is new). *cancelled = 1 when it was still waiting
 * for a lane: it never runs, and its wait returns HDNNP_ERR_CANCELLED (an MD
 * session's later requests still run, from the state without it); 0 when a
 * lane already runs it or it is done, and its wait returns the result. The
 * ticket stays valid until waited for once. */
int hdnnp_server_cancel(hdnnp_server *sv, int64_t ticket, int32_t *cancelled);
/* open an MD session of one system (opened on a lane once it fits) */
int hdnnp_server_md_open(hdnnp_server *sv, int32_t nat, const int32_t *Z, const double *pos, const double *mom,
                         const double *mass, int32_t periodic, const double *lat, double dt, double skin,
                         int64_t *session);
/* queue nsteps velocity-Verlet steps (0: just the energies of the current state) */
int hdnnp_server_md_step(hdnnp_server *sv, int64_t session, int32_t nsteps, int64_t *ticket);
int hdnnp_server_md_wait(hdnnp_server *sv, int64_t ticket, double timeout_s, double *epot, double *ekin,
                         double *latency_ms);
/* state after all the session's queued requests; close frees its device state */
int hdnnp_server_md_get(hdnnp_server *sv, int64_t session, double *pos, double *mom, double *force);
int hdnnp_server_md_close(hdnnp_server *sv, int64_t session);
/* out[16]: evaluations, their atoms, MD requests, MD atom-steps, p50 and
 * p99 evaluation latency (ms), p50 and p99 MD-request latency (ms),
 * batches, mean structures per batch, peak reserved device bytes, budget,
 * seconds since the last reset, most lanes busy at once, refused requests,
 * MD steps. reset != 0 starts a new window. */
int hdnnp_server_stats(hdnnp_server *sv, int32_t reset, double *out);
const char *hdnnp_server_last_error(void);
void hdnnp_server_destroy(hdnnp_server *sv);

/* Replaces dEdLat2Stress(lat, dEdLat, stress) (atomicStructure.f90:679-686):
 * stress = -1/det(lat) * dEdLat . lat^T in the reference's Fortran terms. */
int hdnnp_dEdLat_to_stress(const double *lat, const double *dEdLat, double *stress);

/* Message of the last failed call on ctx (or of the last failed create when
 * ctx is NULL). Never NULL. */
const char *hdnnp_last_error(const hdnnp_ctx *ctx);

/* Timing of the last evaluation's stages in milliseconds (HIP events on the
 * evaluation stream): [0] wrap+cells, [1] neighbour list, [2] symmetry
 * functions, [3] MLP, [4] gradients, [5] force gather + reductions. After an
 * evaluation replayed from a graph (hdnnp_set_graphs), [0] holds its whole
 * device time and the others 0. All zeros when the evaluation ran without
 * stage timings (hdnnp_set_stage_timings). */
int hdnnp_last_timings(hdnnp_ctx *ctx, double *ms6);

/* Stage timings on (1) or off (0, the default; HDNNP_STAGE_TIMINGS=1 in the
 * environment turns them on for every context). No counterpart in the
 * reference. The events between the stages are packets of their own on the
 * stream: they cost an 8-atom evaluation about 12 us and a 2,592-atom one
 * about 20 us. */
int hdnnp_set_stage_timings(hdnnp_ctx *ctx, int32_t on);

/* Graph replay of small evaluations (no counterpart in the reference, whose
 * hdnnpEnergyAndForces, hdnnps.f90:333-371, runs its loops anew on every
 * call). An evaluation of up to 8,192 atoms on the context's own stream
 * (hdnnp_evaluate, hdnnp_evaluate_batch, and the device entry points with
 * stream NULL) is a dozen small kernels. When one has the launch parameters
 * and buffers of the one before it (same atom counts and cells, same
 * neighbour capacity, same input and output arrays for the device entry
 * points), its kernels are captured into a graph; later evaluations with the
 * launch of any of the last four captures replay it: one launch in place of
 * a dozen. Results equal those of direct launches bit for bit. on = 1 or 0
 * (direct launches only, the default unless the environment sets
 * HDNNP_EVAL_GRAPHS=1; drops the captured graphs). On MI355X with ROCm 7 a
 * replay measured slower than the direct launches, whose host side overlaps
 * the kernels: the small evaluations are bound by the kernels' own latency.
 * hdnnp_graph_counts: the mode, and how many evaluations were captured and
 * replayed (each may be NULL). */
int hdnnp_set_graphs(hdnnp_ctx *ctx, int32_t on);
int hdnnp_graph_counts(const hdnnp_ctx *ctx, int32_t *on, int64_t *captures, int64_t *replays);

void hdnnp_destroy(hdnnp_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif
