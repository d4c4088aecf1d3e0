// Cell wrap, cell grid and neighbour-list kernels (FP64, no fused multiply-add:
// this file is built with -ffp-contract=off so that every product and sum is
// rounded exactly as in the reference and the neighbour sets come out the
// same bit for bit).
//
// Reference: hdnnpCalcNeighborLists (src/potentials/hdnnps.f90:373-388),
// moveAtomsIntoCell (src/atoms/pointsets.f90:201-217) and
// buildNeighborListGrid (src/potentials/neighborLists.f90:104-230).
#include "device.hpp"

namespace hdnnp {

#pragma clang fp contract(off)

// matMulVec3 (linalg.f90:194-203): res(r) = sum(M(r,:) * v(:)), M(r,c) = M[r+3c]
__device__ inline void mv3(const double *M, double v0, double v1, double v2, double &o0,
                           double &o1, double &o2)
{
    const double r0 = (M[0] * v0 + M[3] * v1) + M[6] * v2;
    const double r1 = (M[1] * v0 + M[4] * v1) + M[7] * v2;
    const double r2 = (M[2] * v0 + M[5] * v1) + M[8] * v2;
    o0 = r0;
    o1 = r1;
    o2 = r2;
}

// moveAtomsIntoCell for one atom: fractional, modulo 1, Cartesian
__device__ inline void wrap1(const double *lat, const double *inv, double &x, double &y, double &z)
{
    double s0, s1, s2;
    mv3(inv, x, y, z, s0, s1, s2);
    s0 = s0 - floor(s0);
    s1 = s1 - floor(s1);
    s2 = s2 - floor(s2);
    mv3(lat, s0, s1, s2, x, y, z);
}

__device__ inline int floordiv(int a, int b)
{
    int q = a / b;
    if ((a % b != 0) && ((a < 0) != (b < 0))) --q;
    return q;
}

__global__ void k_fill_int(int *p, int v, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

void launch_fill_int(int *p, int v, long n, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill_int, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, v, n);
}

// a[0..na) = 0, b[0..nb) = 0 and c[0..nc) = 0 in one launch
__global__ void k_zero3(int *a, long na, int *b, long nb, int *c, long nc)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < na) a[i] = 0;
    if (i < nb) b[i] = 0;
    if (i < nc) c[i] = 0;
}

void launch_zero3(int *a, long na, int *b, long nb, int *c, long nc, hipStream_t s)
{
    const long n = na > nb ? (na > nc ? na : nc) : (nb > nc ? nb : nc);
    if (n > 0)
        hipLaunchKernelGGL(k_zero3, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a, na, b, nb, c, nc);
}

// Bounding boxes of the raw positions of nonperiodic structures (artificial
// lattice, neighborLists.f90:131-138): block b covers structure which[b],
// atoms [off[s], off[s+1]); out[6b..6b+5] = min x, y, z, max x, y, z.
__global__ __launch_bounds__(256) void k_bbox(const double *pos, const int *off, const int *which, double *out)
{
    __shared__ double red[6][256];
    const int st = which[blockIdx.x];
    const int i0 = off[st], i1 = off[st + 1];
    double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
    for (int i = i0 + threadIdx.x; i < i1; i += blockDim.x)
        for (int k = 0; k < 3; ++k) {
            const double v = pos[3 * (size_t)i + k];
            mn[k] = fmin(mn[k], v);
            mx[k] = fmax(mx[k], v);
        }
    for (int k = 0; k < 3; ++k) {
        red[k][threadIdx.x] = mn[k];
        red[3 + k][threadIdx.x] = mx[k];
    }
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o)
            for (int k = 0; k < 3; ++k) {
                red[k][threadIdx.x] = fmin(red[k][threadIdx.x], red[k][threadIdx.x + o]);
                red[3 + k][threadIdx.x] = fmax(red[3 + k][threadIdx.x], red[3 + k][threadIdx.x + o]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) out[6 * (size_t)blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

void launch_bbox(const double *pos, const int *off, const int *which, int nwhich, double *out, hipStream_t s)
{
    if (nwhich > 0) hipLaunchKernelGGL(k_bbox, dim3(nwhich), dim3(256), 0, s, pos, off, which, out);
}

// inclusive prefix sum over the lanes of a wavefront
__device__ inline int wave_incl_scan_int(int v, int lane)
{
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(v, o);
        if (lane >= o) v += y;
    }
    return v;
}

// The structure of atom x (cells = false) or cell x (cells = true) of a
// batch: the last structure whose first atom / cell is <= x.
__device__ inline int segment_of(const StructGeo *sg, int nstruct, int x, bool cells)
{
    int lo = 0, hi = nstruct - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if ((cells ? sg[mid].cell_off : sg[mid].atom_off) <= x) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// The per-atom set-up of a build in one pass (thread x also covers cell x):
// - the structure of each cell;
// - the element rank of each atom. Unknown Z and elements the model cannot
//   evaluate are flagged (rare, so the atomics are cheap); such atoms get a
//   usable element so that every kernel stays in bounds, and the host
//   discards the results;
// - the structure of each atom and the identity order (the sorts' values);
// - wrap into the list frame and bin: periodic input is wrapped twice (as
//   hdnnps.f90:381 and neighborLists.f90:156 do); nonperiodic input is
//   shifted by the artificial origin and wrapped once in the artificial box.
//   Cell index int(s*ngrid) (neighborLists.f90:161-166), clamped into the
//   grid, offset by the structure's first cell.
// item x of the set-up (cell x, atom x); iota may be null. Returns atom x's
// cell key (-1 past the atoms) and sets el to its element rank.
__device__ __forceinline__ int prep_one(const DevModel &m, const Geo &g, const Work &w, int *iota, int x, int &el)
{
    if (x < g.ncell) w.csid[x] = g.nstruct == 1 ? 0 : segment_of(g.sg, g.nstruct, x, true);
    if (x >= w.nat) return -1;
    const int z = w.Z[x];
    const int e = (z >= 0 && z < 128) ? m.zlut[z] : -1;
    if (e == -1) atomicMax(&w.flags[0], z > 0 ? z : 1);
    if (e <= -2) atomicMax(&w.flags[3], z);
    el = e < 0 ? m.fallback_el : e;
    w.eidx[x] = el;
    const int sid = g.nstruct == 1 ? 0 : segment_of(g.sg, g.nstruct, x, false);
    w.sid[x] = sid;
    if (iota) iota[x] = x;
    const StructGeo &G = g.sg[sid];
    double px = w.pos[3 * x + 0], py = w.pos[3 * x + 1], pz = w.pos[3 * x + 2];
    if (G.periodic) {
        wrap1(G.lat, G.inv, px, py, pz);
        wrap1(G.lat, G.inv, px, py, pz);
    } else {
        px = px - G.origin[0];
        py = py - G.origin[1];
        pz = pz - G.origin[2];
        wrap1(G.lat, G.inv, px, py, pz);
    }
    w.wpos[x] = make_double4(px, py, pz, 0.0);
    double s0, s1, s2;
    mv3(G.inv, px, py, pz, s0, s1, s2);
    int c0 = (int)(s0 * G.ng[0]), c1 = (int)(s1 * G.ng[1]), c2 = (int)(s2 * G.ng[2]);
    c0 = min(max(c0, 0), G.ng[0] - 1);
    c1 = min(max(c1, 0), G.ng[1] - 1);
    c2 = min(max(c2, 0), G.ng[2] - 1);
    const int key = G.cell_off + (c0 * G.ng[1] + c1) * G.ng[2] + c2;
    w.cellkey[x] = key;
    return key;
}

// slot (or null): atom x's arrival slot in its cell, counted in cell_end;
// chunk_cnt (or null): the atoms of each element in each kSortChunk-atom
// chunk, the element sort's counts (both zero before the launch)
constexpr int kSortChunk = 1024;   // atoms per chunk of the element sort: 4 rounds of 256 threads
__global__ __launch_bounds__(256) void k_prep(DevModel m, Geo g, Work w, int *iota, int *slot, int *chunk_cnt)
{
    __shared__ int ecnt[kMaxEl];
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (chunk_cnt && threadIdx.x < kMaxEl) ecnt[threadIdx.x] = 0;
    int el = -1;
    const int key = prep_one(m, g, w, iota, x, el);
    if (slot && key >= 0) slot[x] = atomicAdd(&w.cell_end[key], 1);
    if (chunk_cnt) {   // this block's 256 atoms lie in one chunk
        __syncthreads();
        if (key >= 0) atomicAdd(&ecnt[el], 1);
        __syncthreads();
        const int c = (int)(blockIdx.x * blockDim.x) / kSortChunk;
        if (threadIdx.x < kMaxEl && ecnt[threadIdx.x])
            atomicAdd(&chunk_cnt[c * kMaxEl + threadIdx.x], ecnt[threadIdx.x]);
    }
}

void launch_prep(const DevModel &m, const Geo &g, Work &w, int *iota, int *slot, int *chunk_cnt, hipStream_t s)
{
    const int n = w.nat > g.ncell ? w.nat : g.ncell;
    if (n > 0)
        hipLaunchKernelGGL(k_prep, dim3((n + 255) / 256), dim3(256), 0, s, m, g, w, iota, slot, chunk_cnt);
}

// Counting sort by cell, for grids above the one-workgroup sort's size: after
// k_prep has counted each cell's atoms in cell_end and kept each atom's
// arrival slot, k_cell_scan (one workgroup) turns the counts into each
// cell's range, k_cell_place puts atom x at its arrival slot of its cell in
// tmp, and k_cell_rank moves it to its rank in index order (the number of the
// cell's atoms with a smaller index), which is the stable radix sort's order.
// Three launches where the radix sort and the range search take eight.
// Cells in rows of 1024, thread t taking column t (every access coalesced),
// kScanRows rows per round (kCountSortCells cells: one round), their loads
// all out at once: each row's wave scans, then one wave scans the
// (row, wave) totals in cell order.
constexpr int kScanRows = 24;
static_assert(kCountSortCells == kScanRows * 1024, "the counting sort's grids take one round of the scan");
__device__ void elem_offsets_block(Work &w, int *chunk_cnt, int nchunk);
// then, as k_elem_offsets, the element sort's chunk offsets from chunk_cnt
__global__ __launch_bounds__(1024) void k_cell_scan(Work w, int ncell, int *chunk_cnt, int nchunk)
{
    __shared__ int pre[kScanRows * 16];
    __shared__ int round_total;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    int carry = 0;   // the cells of the rounds before
    for (int r0 = 0; r0 * 1024 < ncell; r0 += kScanRows) {
        int v[kScanRows], inc[kScanRows];
#pragma unroll
        for (int u = 0; u < kScanRows; ++u) {
            const int c = (r0 + u) * 1024 + t;
            v[u] = c < ncell ? w.cell_end[c] : 0;
        }
#pragma unroll
        for (int u = 0; u < kScanRows; ++u) {
            inc[u] = wave_incl_scan_int(v[u], lane);
            if (lane == 63) pre[u * 16 + wv] = inc[u];
        }
        __syncthreads();
        if (wv == 0) {   // exclusive scan of the kScanRows * 16 totals, kPerLane per lane
            constexpr int kPerLane = kScanRows * 16 / 64;
            static_assert(kScanRows % 4 == 0, "whole totals per lane");
            int a[kPerLane], sum = 0;
#pragma unroll
            for (int i = 0; i < kPerLane; ++i) {
                a[i] = pre[lane * kPerLane + i];
                sum += a[i];
            }
            const int ps = wave_incl_scan_int(sum, lane);
            int run = ps - sum;
#pragma unroll
            for (int i = 0; i < kPerLane; ++i) {
                pre[lane * kPerLane + i] = run;
                run += a[i];
            }
            if (lane == 63) round_total = ps;
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < kScanRows; ++u) {
            const int c = (r0 + u) * 1024 + t;
            const int st = carry + pre[u * 16 + wv] + inc[u] - v[u];
            if (c < ncell) {
                w.cell_start[c] = st;
                w.cell_end[c] = st + v[u];
            }
        }
        carry += round_total;
        __syncthreads();   // pre and round_total are read before the next round writes them
    }
    elem_offsets_block(w, chunk_cnt, nchunk);
}

__device__ void elem_place_block(Work &w, const int *chunk_off, int nel, int chunk);

// blocks [0, nplace): atom x to its arrival slot of its cell in tmp; the
// blocks after: the element sort's placement, chunk by chunk (the two are
// independent, so they share a launch)
__global__ __launch_bounds__(256) void k_cell_place(Work w, const int *slot, int *tmp, const int *chunk_off, int nel,
                                                    int nplace)
{
    if ((int)blockIdx.x >= nplace) {   // block-uniform
        elem_place_block(w, chunk_off, nel, (int)blockIdx.x - nplace);
        return;
    }
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x < w.nat) tmp[w.cell_start[w.cellkey[x]] + slot[x]] = x;
}

__global__ __launch_bounds__(256) void k_cell_rank(Work w, const int *tmp)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= w.nat) return;
    const int c = w.cellkey[x], s0 = w.cell_start[c], s1 = w.cell_end[c];
    int r = 0;
    for (int q = s0; q < s1; ++q) r += tmp[q] < x;
    w.cell_atoms[s0 + r] = x;
}

void launch_cell_count_sort(Work &w, int ncell, int nel, const int *slot, int *tmp, int *chunk_cnt, hipStream_t s)
{
    if (w.nat <= 0) return;
    hipLaunchKernelGGL(k_cell_scan, dim3(1), dim3(1024), 0, s, w, ncell, chunk_cnt, elem_sort_chunks(w.nat));
    const int nplace = (w.nat + 255) / 256;
    hipLaunchKernelGGL(k_cell_place, dim3(nplace + elem_sort_chunks(w.nat)), dim3(256), 0, s, w, slot, tmp,
                       (const int *)chunk_cnt, nel, nplace);
    hipLaunchKernelGGL(k_cell_rank, dim3((w.nat + 255) / 256), dim3(256), 0, s, w, (const int *)tmp);
}

// first position q in [0, n) with keys[q] >= v
__device__ inline int lower_bound(const int *keys, int n, int v)
{
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (keys[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// The atoms of cell c are cell_atoms[cell_start[c] .. cell_end[c]) of the
// key-sorted order (an empty cell has start = end).
__global__ __launch_bounds__(256) void k_cell_bounds(Work w, int ncell)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    w.cell_start[c] = lower_bound(w.cellkey_sorted, w.nat, c);
    w.cell_end[c] = lower_bound(w.cellkey_sorted, w.nat, c + 1);
}

void launch_cell_bounds(Work &w, int ncell, hipStream_t s)
{
    if (ncell > 0) hipLaunchKernelGGL(k_cell_bounds, dim3((ncell + 255) / 256), dim3(256), 0, s, w, ncell);
}

// The sorts of a small evaluation in one workgroup (one launch where the
// radix sort, the range search and the element sort take eleven).
// By cell: per-cell counts in LDS, their exclusive scan, placement by arrival
// slot, then the order within a cell by atom index, which is the stable
// sort's order (atom i's rank is the number of its cell's atoms with a
// smaller index). By element: counts, offsets, then placement in rounds of
// the workgroup's size in index order, ranked by wave ballots as in
// k_elem_place. The per-atom set-up of k_prep runs first, in the same
// workgroup. Needs nat <= kSmallSortAtoms and ncell <= kSmallSortCells;
// slot and tmp hold nat ints each.
// Ranking an atom within its cell by a scan of the cell costs k loads for a
// cell of k atoms, k^2 for the cell. When the fullest cell holds more than
// kScanRankMax atoms, the atoms are placed instead in rounds of the
// workgroup's size in index order: an atom's rank among the round's atoms of
// its cell comes from a scan of the round's cell keys in LDS, and the first
// of them takes the round's positions of its cell from a running counter
// (cell_end, which ends at the cell's end). That costs nat loads from LDS per
// thread whatever the distribution.
constexpr int kScanRankMax = 256;
__global__ __launch_bounds__(kSmallSortThreads) void k_sort_small(DevModel m, Geo g, Work w, int nel, int *slot,
                                                                  int *tmp, int *clear, int nclear)
{
    __shared__ int start[kSmallSortCells + 1];
    __shared__ int part[kSmallSortThreads];
    __shared__ int fullest;
    const int t = threadIdx.x, nat = w.nat, ncell = g.ncell;
    // the flags and counters of the evaluation start at zero (in place of a
    // fill launch): cleared before any thread of this workgroup, the only
    // one, sets a flag; the barrier's fence has the stores reach L2, where
    // the flag atomics act
    for (int x = t; x < nclear; x += kSmallSortThreads) clear[x] = 0;
    __syncthreads();
    for (int x = t; x < max(nat, ncell); x += kSmallSortThreads) {
        int el;
        prep_one(m, g, w, nullptr, x, el);
    }
    if (t == 0) fullest = 0;
    __syncthreads();
    for (int c = t; c <= ncell; c += kSmallSortThreads) start[c] = 0;
    __syncthreads();
    for (int i = t; i < nat; i += kSmallSortThreads) slot[i] = atomicAdd(&start[w.cellkey[i]], 1);
    __syncthreads();
    // exclusive scan of the counts: thread t owns cells [c0, c1)
    const int per = (ncell + kSmallSortThreads - 1) / kSmallSortThreads;
    const int c0 = min(t * per, ncell), c1 = min(c0 + per, ncell);
    int sum = 0, most = 0;
    for (int c = c0; c < c1; ++c) {
        sum += start[c];
        most = max(most, start[c]);
    }
    atomicMax(&fullest, most);
    // the sums of the threads before: wave scans, then the totals of the
    // waves before (one barrier)
    __shared__ int wtot[kSmallSortThreads / 64];
    const int incl = wave_incl_scan_int(sum, t & 63);
    if ((t & 63) == 63) wtot[t >> 6] = incl;
    __syncthreads();
    int run = incl - sum;
    for (int k = 0; k < (t >> 6); ++k) run += wtot[k];
    const bool by_scan = fullest <= kScanRankMax;   // workgroup-uniform
    for (int c = c0; c < c1; ++c) {
        const int n = start[c];
        start[c] = run;
        w.cell_start[c] = run;
        w.cell_end[c] = by_scan ? run + n : run;   // else a running counter from the cell's start
        run += n;
    }
    if (t == 0) start[ncell] = nat;
    // the running counters are advanced by atomics, which act in L2: the
    // stores above must have reached it before any thread adds to them
    if (!by_scan) __threadfence();
    __syncthreads();
    if (by_scan) {
        for (int i = t; i < nat; i += kSmallSortThreads) tmp[start[w.cellkey[i]] + slot[i]] = i;
        __syncthreads();
        for (int i = t; i < nat; i += kSmallSortThreads) {
            const int c = w.cellkey[i], s0 = start[c], s1 = start[c + 1];
            int r = 0;
            for (int q = s0; q < s1; ++q) r += tmp[q] < i;
            w.cell_atoms[s0 + r] = i;
        }
    } else {
        __shared__ int rbase[kSmallSortThreads];   // at a cell's first atom of the round: its base
        int *rkey = part;                           // the round's cell keys
        for (int r0 = 0; r0 < nat; r0 += kSmallSortThreads) {
            const int i = r0 + t, n = min(kSmallSortThreads, nat - r0);
            const int c = i < nat ? w.cellkey[i] : -1;
            __syncthreads();                 // the previous round has read rkey
            rkey[t] = c;
            __syncthreads();
            int rank = 0, tot = 0, first = t;
            if (c >= 0)
                for (int j = 0; j < n; ++j) {
                    const bool same = rkey[j] == c;
                    tot += same;
                    if (same && j < t) {
                        ++rank;
                        first = min(first, j);
                    }
                }
            if (c >= 0 && rank == 0) rbase[t] = atomicAdd(&w.cell_end[c], tot);
            __syncthreads();
            if (c >= 0) w.cell_atoms[rbase[first] + rank] = i;
        }
    }

    // by element (the MLP tiles)
    constexpr int kWaves = kSmallSortThreads / 64;
    __shared__ int ecnt[kMaxEl], ebase[kMaxEl], wcnt[kWaves][kMaxEl];
    const int lane = t % 64, wave = t / 64;
    if (t < kMaxEl) ecnt[t] = 0;
    __syncthreads();
    for (int i = t; i < nat; i += kSmallSortThreads) atomicAdd(&ecnt[w.eidx[i]], 1);
    __syncthreads();
    if (t <= kMaxEl) {
        int off = 0;
        for (int f = 0; f < t; ++f) off += ecnt[f];
        w.elem_off[t] = off;
        if (t < kMaxEl) {
            w.elem_count[t] = ecnt[t];
            ebase[t] = off;
        }
    }
    __syncthreads();
    const unsigned long long below = (1ull << lane) - 1;
    for (int r0 = 0; r0 < nat; r0 += kSmallSortThreads) {
        const int i = r0 + t;
        const int e = i < nat ? w.eidx[i] : -1;
        int rank = 0;
        for (int f = 0; f < nel; ++f) {
            const unsigned long long mk = __ballot(e == f);
            if (e == f) rank = __popcll(mk & below);
            if (lane == 0) wcnt[wave][f] = __popcll(mk);
        }
        __syncthreads();                           // wcnt of this round, ebase of the rounds before
        if (e >= 0) {
            int pos = ebase[e] + rank;
            for (int v = 0; v < wave; ++v) pos += wcnt[v][e];
            w.perm[pos] = i;
        }
        __syncthreads();                           // every thread has read ebase and wcnt
        if (t < nel) {
            int add = 0;
            for (int v = 0; v < kWaves; ++v) add += wcnt[v][t];
            ebase[t] += add;
        }
        __syncthreads();                           // ebase updated before the next round's wcnt
    }
}

void launch_sort_small(const DevModel &m, const Geo &g, Work &w, int *slot, int *tmp, int *clear, int nclear,
                       hipStream_t s)
{
    hipLaunchKernelGGL(k_sort_small, dim3(1), dim3(kSmallSortThreads), 0, s, m, g, w, m.nel, slot, tmp, clear,
                       nclear);
}

// Atoms ordered by element, in index order within an element (the tiles of
// the MLP), by a stable counting sort over chunks of kSortChunk atoms:
// per-chunk counts, their offsets, then each chunk places its atoms.
__global__ __launch_bounds__(256) void k_elem_count(Work w, int *chunk_cnt)
{
    __shared__ int cnt[kMaxEl];
    if (threadIdx.x < kMaxEl) cnt[threadIdx.x] = 0;
    __syncthreads();
    const int a0 = blockIdx.x * kSortChunk;
    for (int r = 0; r < kSortChunk; r += 256) {
        const int i = a0 + r + (int)threadIdx.x;
        if (i < w.nat) atomicAdd(&cnt[w.eidx[i]], 1);
    }
    __syncthreads();
    if (threadIdx.x < kMaxEl) chunk_cnt[blockIdx.x * kMaxEl + threadIdx.x] = cnt[threadIdx.x];
}

// one block of kMaxEl wavefronts: wave e scans element e's chunk counts into
// chunk offsets; then elem_count[e], and elem_off[e] = first position of
// element e (elem_off[kMaxEl] = nat)
static_assert(64 * kMaxEl == 1024, "the element offsets run in the cell scan's workgroup");
__device__ void elem_offsets_block(Work &w, int *chunk_cnt, int nchunk)
{
    __shared__ int tot[kMaxEl];
    const int e = threadIdx.x / 64, lane = threadIdx.x % 64;
    int base = 0;
    for (int c0 = 0; c0 < nchunk; c0 += 64) {
        const int c = c0 + lane;
        const int v = c < nchunk ? chunk_cnt[c * kMaxEl + e] : 0;
        const int incl = wave_incl_scan_int(v, lane);
        if (c < nchunk) chunk_cnt[c * kMaxEl + e] = base + incl - v;
        base += __shfl(incl, 63);
    }
    if (lane == 0) tot[e] = base;
    __syncthreads();
    if (threadIdx.x <= kMaxEl) {
        int off = 0;
        for (int f = 0; f < (int)threadIdx.x; ++f) off += tot[f];
        w.elem_off[threadIdx.x] = off;
        if (threadIdx.x < kMaxEl) w.elem_count[threadIdx.x] = tot[threadIdx.x];
    }
}

__global__ __launch_bounds__(64 * kMaxEl) void k_elem_offsets(Work w, int *chunk_cnt, int nchunk)
{
    elem_offsets_block(w, chunk_cnt, nchunk);
}

__device__ void elem_place_block(Work &w, const int *chunk_off, int nel, int chunk)
{
    __shared__ int wcnt[4][kMaxEl];
    __shared__ int base[kMaxEl];
    const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    if (threadIdx.x < kMaxEl) base[threadIdx.x] = w.elem_off[threadIdx.x] + chunk_off[chunk * kMaxEl + threadIdx.x];
    const unsigned long long below = (1ull << lane) - 1;
    const int a0 = chunk * kSortChunk;
    for (int r = 0; r < kSortChunk; r += 256) {
        const int i = a0 + r + (int)threadIdx.x;
        const int e = i < w.nat ? w.eidx[i] : -1;
        int rank = 0;
        for (int f = 0; f < nel; ++f) {
            const unsigned long long mk = __ballot(e == f);
            if (e == f) rank = __popcll(mk & below);
            if (lane == 0) wcnt[wave][f] = __popcll(mk);
        }
        __syncthreads();                           // wcnt of this round, base of the rounds before
        if (e >= 0) {
            int pos = base[e] + rank;
            for (int v = 0; v < wave; ++v) pos += wcnt[v][e];
            w.perm[pos] = i;
        }
        __syncthreads();                           // every lane has read base and wcnt
        if (threadIdx.x < nel) base[threadIdx.x] += wcnt[0][threadIdx.x] + wcnt[1][threadIdx.x] +
                                                    wcnt[2][threadIdx.x] + wcnt[3][threadIdx.x];
        __syncthreads();                           // base updated before the next round's wcnt
    }
}

__global__ __launch_bounds__(256) void k_elem_place(Work w, const int *chunk_off, int nel)
{
    elem_place_block(w, chunk_off, nel, (int)blockIdx.x);
}

int elem_sort_chunks(int nat) { return (nat + kSortChunk - 1) / kSortChunk; }

void launch_elem_sort(Work &w, int nel, int *chunk_cnt, hipStream_t s)
{
    const int nchunk = elem_sort_chunks(w.nat);
    if (nchunk > 0) hipLaunchKernelGGL(k_elem_count, dim3(nchunk), dim3(256), 0, s, w, chunk_cnt);
    hipLaunchKernelGGL(k_elem_offsets, dim3(1), dim3(64 * kMaxEl), 0, s, w, chunk_cnt, nchunk);
    if (nchunk > 0) hipLaunchKernelGGL(k_elem_place, dim3(nchunk), dim3(256), 0, s, w, chunk_cnt, nel);
}

// One stencil cell of atom i: wrapped cell id (within the structure's grid)
// and image shift, in the reference's order of wrapping (z, then y, then x;
// neighborLists.f90:186-209). Returns false when the cell lies outside a
// nonperiodic grid.
struct StencilCell {
    int cell;
    int sx, sy, sz;
    double dl0, dl1, dl2;
    bool centre;
};

__device__ inline bool stencil_cell(const StructGeo &g, int key, int s, StencilCell &c)
{
    const int sz = 2 * g.nc[2] + 1, sy = 2 * g.nc[1] + 1;
    const int jz = s % sz - g.nc[2];
    const int t = s / sz;
    const int jy = t % sy - g.nc[1];
    const int jx = t / sy - g.nc[0];
    const int iz = key % g.ng[2], iy = (key / g.ng[2]) % g.ng[1], ix = key / (g.ng[1] * g.ng[2]);
    int cx = ix + jx, cy = iy + jy, cz = iz + jz;
    c.sx = c.sy = c.sz = 0;
    c.dl0 = c.dl1 = c.dl2 = 0.0;
    c.centre = (jx == 0 && jy == 0 && jz == 0);
    if (cz < 0 || cz >= g.ng[2]) {
        if (!g.periodic) return false;
        const int d = floordiv(cz, g.ng[2]);
        cz -= g.ng[2] * d;
        c.dl0 = c.dl0 + g.lat[6] * (double)d;
        c.dl1 = c.dl1 + g.lat[7] * (double)d;
        c.dl2 = c.dl2 + g.lat[8] * (double)d;
        c.sz = d;
    }
    if (cy < 0 || cy >= g.ng[1]) {
        if (!g.periodic) return false;
        const int d = floordiv(cy, g.ng[1]);
        cy -= g.ng[1] * d;
        c.dl0 = c.dl0 + g.lat[3] * (double)d;
        c.dl1 = c.dl1 + g.lat[4] * (double)d;
        c.dl2 = c.dl2 + g.lat[5] * (double)d;
        c.sy = d;
    }
    if (cx < 0 || cx >= g.ng[0]) {
        if (!g.periodic) return false;
        const int d = floordiv(cx, g.ng[0]);
        cx -= g.ng[0] * d;
        c.dl0 = c.dl0 + g.lat[0] * (double)d;
        c.dl1 = c.dl1 + g.lat[1] * (double)d;
        c.dl2 = c.dl2 + g.lat[2] * (double)d;
        c.sx = d;
    }
    c.cell = (cx * g.ng[1] + cy) * g.ng[2] + cz;
    return true;
}

// Neighbour lists of the atoms of one grid cell per wave (one-wave blocks),
// or, on grids of few cells, of every split-th atom of the cell (wave p
// takes atoms p, p + split, ...; the waves of a cell run on one XCD,
// xcd_block): more waves in flight where the cells alone leave SIMDs idle.
// The atoms of a wave share the stencil, so the candidates are enumerated
// once and tested against each of them.
//
// Per chunk of 64 stencil cells each lane resolves one cell (wrapped id,
// image shift and the lattice offset dl, summed as the reference does); the
// chunk's candidate atoms are numbered in stencil order, atoms of a cell in
// ascending index, and read 64 at a time: lane t finds its cell by a binary
// search over the chunk's prefix counts. For each atom of the cell a ballot
// compacts the hits, so its entries are written in exactly the reference's
// list order. pos_start (when present) receives, per atom and stencil cell,
// the slot of that cell's first entry; the partner search uses it.
// Entries beyond `cap` are counted but not stored; flags[1] then receives
// the number needed so that the host can repeat the evaluation.
__global__ __launch_bounds__(64) void k_nl_build(Geo g, Work w, int split)
{
    __shared__ int pref[64], q0[64], info[64], ia[64];
    __shared__ double dl[3][64];
    __shared__ double4 pa[64];
    const int lane = threadIdx.x;
    const int lb = split > 1 ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
    const int c = lb / split, part = lb - c * split;
    const int a0 = w.cell_start[c], A = (w.cell_end[c] - a0 - part + split - 1) / split;   // this wave's atoms
    if (A <= 0) return;
    const StructGeo &G = g.sg[w.csid[c]];
    const int S = G.S;
    const int key = c - G.cell_off;
    for (int g0 = 0; g0 < A; g0 += 64) {
        const int Ag = min(64, A - g0);
        __syncthreads();                      // previous group done with ia / pa
        if (lane < Ag) {
            const int t = w.cell_atoms[a0 + part + split * (g0 + lane)];
            ia[lane] = t;
            pa[lane] = w.wpos[t];
        }
        int n = 0;                            // lane a: entries of atom a so far
        for (int s0 = 0; s0 < S; s0 += 64) {
            StencilCell sc{};
            int m = 0, start = 0;
            const bool in_s = s0 + lane < S;
            if (in_s && stencil_cell(G, key, s0 + lane, sc)) {
                start = w.cell_start[G.cell_off + sc.cell];
                m = w.cell_end[G.cell_off + sc.cell] - start;
            }
            const int incl = wave_incl_scan_int(m, lane);
            const int tot = __shfl(incl, 63);
            const int my_pref = incl - m;
            __syncthreads();                  // previous chunk done with the tables
            pref[lane] = my_pref;
            q0[lane] = start;
            const bool big = sc.sx < -127 || sc.sx > 127 || sc.sy < -127 || sc.sy > 127 || sc.sz < -127 ||
                             sc.sz > 127;
            info[lane] = m ? (pack_shift(sc.sx, sc.sy, sc.sz) | (sc.centre ? 1 << 24 : 0) | (big ? 1 << 25 : 0)) : 0;
            dl[0][lane] = sc.dl0;
            dl[1][lane] = sc.dl1;
            dl[2][lane] = sc.dl2;
            __syncthreads();
            for (int t0 = 0; t0 < tot; t0 += 64) {
                const int t = t0 + lane;
                int j = -1, o = 0, inf = 0;
                double e0 = 0, e1 = 0, e2 = 0;
                if (t < tot) {
#pragma unroll
                    for (int b = 32; b > 0; b >>= 1)
                        if (pref[o + b] <= t) o += b;
                    j = w.cell_atoms[q0[o] + (t - pref[o])];
                    inf = info[o];
                    const double4 pj = w.wpos[j];
                    e0 = pj.x + dl[0][o];             // image of j: pj + dl, as neighborLists.f90:212
                    e1 = pj.y + dl[1][o];
                    e2 = pj.z + dl[2][o];
                }
                // this stencil lane's first candidate falls in this round
                const bool first_here = in_s && my_pref >= t0 && my_pref < t0 + 64;
                for (int a = 0; a < Ag; ++a) {
                    const int ai = ia[a];
                    bool hit = false;
                    if (t < tot && !(j == ai && (inf >> 24 & 1))) {
                        const double4 p = pa[a];
                        const double d0 = p.x - e0, d1 = p.y - e1, d2 = p.z - e2;
                        hit = (d0 * d0 + d1 * d1) + d2 * d2 < g.rc2;    // neighborLists.f90:213-214
                    }
                    const unsigned long long mask = __ballot(hit);
                    const int na = __shfl(n, a);
                    if (hit) {
                        const int pos = na + (int)__builtin_amdgcn_mbcnt_hi(
                                                 (unsigned)(mask >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0));
                        if (inf >> 25 & 1) atomicMax(&w.flags[2], 1);
                        if (pos < w.lim) {
                            const int q = ai * w.cap + pos;
                            w.ent_j[q] = j;
                            w.ent_shift[q] = inf & 0xFFFFFF;
                            w.ent_s[q] = s0 + o;
                        }
                    }
                    if (w.pos_start && first_here) {
                        const int k = my_pref - t0;
                        const int ps = na + __popcll(mask & ((1ull << k) - 1));
                        w.pos_start[G.ps_off + (long long)(ai - G.atom_off) * S + s0 + lane] =
                            (unsigned short)min(ps, 65535);
                    }
                    if (lane == a) n += __popcll(mask);
                }
            }
            // stencil cells of this chunk that start at its end (empty tail cells)
            if (w.pos_start)
                for (int a = 0; a < Ag; ++a) {
                    const int na = __shfl(n, a);
                    if (in_s && my_pref >= tot)
                        w.pos_start[G.ps_off + (long long)(ia[a] - G.atom_off) * S + s0 + lane] =
                            (unsigned short)min(na, 65535);
                }
        }
        if (lane < Ag) {
            w.count[ia[lane]] = min(n, w.lim);
            if (n > w.lim) atomicMax(&w.flags[1], n);
        }
    }
}

void launch_nl_build(const Geo &g, Work &w, int ncell, hipStream_t s)
{
    // waves per cell: up to 4 while the grid has fewer than 8,192 cells
    const int split = ncell >= 8192 ? 1 : ncell >= 4096 ? 2 : 4;
    if (ncell > 0) hipLaunchKernelGGL(k_nl_build, dim3(ncell * split), dim3(64), 0, s, g, w, split);
}

// Partner entries: entry (i -> j, image L) found in stencil cell s of i pairs
// with (j -> i, -L), which j finds in the mirrored stencil cell S-1-s. With
// pos_start the partner lies among the few entries of that cell in j's list;
// without it, lists are sorted by (stencil cell, atom) and are searched
// by bisection. The strict distance test is evaluated separately in each
// direction, so a pair at the exact cutoff can lack its partner: rev = -1.
// One wavefront per atom (four per block), its lanes over the atom's entries.
__global__ __launch_bounds__(256) void k_nl_rev(Geo g, Work w)
{
    const int i = xcd_block(blockIdx.x, gridDim.x) * 4 + (int)threadIdx.x / 64;
    if (i >= w.nat) return;
    const int cnt = w.count[i];
    const StructGeo &G = g.sg[w.sid[i]];
    const int S = G.S;
    for (int t = (int)threadIdx.x % 64; t < cnt; t += 64) {
        const long q = (long)i * w.cap + t;
        const int j = w.ent_j[q];
        const int sp = S - 1 - w.ent_s[q];
        const int bj = j * w.cap, nj = w.count[j];
        int r = -1;
        if (w.pos_start) {
            for (int u = w.pos_start[G.ps_off + (long long)(j - G.atom_off) * S + sp]; u < nj; ++u) {
                if (w.ent_s[bj + u] != sp) break;
                if (w.ent_j[bj + u] == i) {
                    r = bj + u;
                    break;
                }
            }
        } else {
            int lo = 0, hi = nj;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                const int st = w.ent_s[bj + mid];
                if (st < sp || (st == sp && w.ent_j[bj + mid] < i)) lo = mid + 1;
                else hi = mid;
            }
            if (lo < nj && w.ent_s[bj + lo] == sp && w.ent_j[bj + lo] == i) r = bj + lo;
        }
        w.rev[q] = r;
    }
}

void launch_nl_rev(const Geo &g, Work &w, hipStream_t s)
{
    if (w.nat > 0) hipLaunchKernelGGL(k_nl_rev, dim3((w.nat + 3) / 4), dim3(256), 0, s, g, w);
}

}  // namespace hdnnp
