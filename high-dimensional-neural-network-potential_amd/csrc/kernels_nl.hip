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

// Element rank of each atom. Unknown Z and elements the model cannot
// evaluate are flagged (rare, so the atomics are cheap); such atoms get a
// usable element so that every kernel stays in bounds, and the host discards
// the results.
__global__ __launch_bounds__(256) void k_map_elements(DevModel m, Work w)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= w.nat) return;
    const int z = w.Z[i];
    const int e = (z >= 0 && z < 128) ? m.zlut[z] : -1;
    if (e == -1) atomicMax(&w.flags[0], z > 0 ? z : 1);
    if (e <= -2) atomicMax(&w.flags[3], z);
    w.eidx[i] = e < 0 ? m.fallback_el : e;
}

void launch_map_elements(const DevModel &m, Work &w, hipStream_t s)
{
    hipLaunchKernelGGL(k_map_elements, dim3((w.nat + 255) / 256), dim3(256), 0, s, m, w);
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

// Structure of each atom (cells = false) or each cell (cells = true) of a
// batch: the last structure whose first atom / cell is <= x.
__global__ __launch_bounds__(256) void k_segment_ids(int *out, int n, const StructGeo *sg, int nstruct, bool cells)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= n) return;
    int lo = 0, hi = nstruct - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if ((cells ? sg[mid].cell_off : sg[mid].atom_off) <= x) lo = mid;
        else hi = mid - 1;
    }
    out[x] = lo;
}

void launch_segment_ids(int *out, int n, const StructGeo *sg, int nstruct, bool cells, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_segment_ids, dim3((n + 255) / 256), dim3(256), 0, s, out, n, sg, nstruct, cells);
}

// Wrap into the list frame and bin: periodic input is wrapped twice (as
// hdnnps.f90:381 and neighborLists.f90:156 do); nonperiodic input is shifted
// by the artificial origin and wrapped once in the artificial box. Cell index
// int(s*ngrid) (neighborLists.f90:161-166), clamped into the grid, offset by
// the structure's first cell.
__global__ __launch_bounds__(256) void k_wrap_bin(Geo g, Work w)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= w.nat) return;
    const StructGeo &G = g.sg[w.sid[i]];
    double x = w.pos[3 * i + 0], y = w.pos[3 * i + 1], z = w.pos[3 * i + 2];
    if (G.periodic) {
        wrap1(G.lat, G.inv, x, y, z);
        wrap1(G.lat, G.inv, x, y, z);
    } else {
        x = x - G.origin[0];
        y = y - G.origin[1];
        z = z - G.origin[2];
        wrap1(G.lat, G.inv, x, y, z);
    }
    w.wpos[i] = make_double4(x, y, z, 0.0);
    double s0, s1, s2;
    mv3(G.inv, x, y, z, s0, s1, s2);
    int c0 = (int)(s0 * G.ng[0]), c1 = (int)(s1 * G.ng[1]), c2 = (int)(s2 * G.ng[2]);
    c0 = min(max(c0, 0), G.ng[0] - 1);
    c1 = min(max(c1, 0), G.ng[1] - 1);
    c2 = min(max(c2, 0), G.ng[2] - 1);
    w.cellkey[i] = G.cell_off + (c0 * G.ng[1] + c1) * G.ng[2] + c2;
}

void launch_wrap_bin(const Geo &g, Work &w, hipStream_t s)
{
    hipLaunchKernelGGL(k_wrap_bin, dim3((w.nat + 255) / 256), dim3(256), 0, s, g, w);
}

__global__ __launch_bounds__(256) void k_cell_range(Work w)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= w.nat) return;
    const int key = w.cellkey_sorted[q];
    if (q == 0 || w.cellkey_sorted[q - 1] != key) w.cell_start[key] = q;
    if (q == w.nat - 1 || w.cellkey_sorted[q + 1] != key) w.cell_end[key] = q + 1;
}

void launch_cell_range(const Geo &g, Work &w, int ncell, hipStream_t s)
{
    launch_fill_int(w.cell_start, 0, ncell, s);
    launch_fill_int(w.cell_end, 0, ncell, s);
    hipLaunchKernelGGL(k_cell_range, dim3((w.nat + 255) / 256), dim3(256), 0, s, w);
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

__device__ inline int wave_incl_scan_int(int v, int lane)
{
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(v, o);
        if (lane >= o) v += y;
    }
    return v;
}

// Neighbour lists of the atoms of one grid cell per wave (one-wave blocks).
// All atoms of a cell share the stencil, so the candidates are enumerated
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
__global__ __launch_bounds__(64) void k_nl_build(Geo g, Work w)
{
    __shared__ int pref[64], q0[64], info[64], ia[64];
    __shared__ double dl[3][64];
    __shared__ double4 pa[64];
    const int lane = threadIdx.x;
    const int c = blockIdx.x;
    const int a0 = w.cell_start[c], A = w.cell_end[c] - a0;
    if (A <= 0) return;
    const StructGeo &G = g.sg[w.csid[c]];
    const int S = G.S;
    const int key = c - G.cell_off;
    for (int g0 = 0; g0 < A; g0 += 64) {
        const int Ag = min(64, A - g0);
        __syncthreads();                      // previous group done with ia / pa
        if (lane < Ag) {
            const int t = w.cell_atoms[a0 + g0 + lane];
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
                        if (pos < w.cap) {
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
            w.count[ia[lane]] = min(n, w.cap);
            if (n > w.cap) atomicMax(&w.flags[1], n);
        }
    }
}

void launch_nl_build(const Geo &g, Work &w, int ncell, hipStream_t s)
{
    hipLaunchKernelGGL(k_nl_build, dim3(ncell), dim3(64), 0, s, g, w);
}

// Partner entries: entry (i -> j, image L) found in stencil cell s of i pairs
// with (j -> i, -L), which j finds in the mirrored stencil cell S-1-s. With
// pos_start the partner lies among the few entries of that cell in j's list;
// without it, lists are sorted by (stencil cell, atom) and are searched
// by bisection. The strict distance test is evaluated separately in each
// direction, so a pair at the exact cutoff can lack its partner: rev = -1.
// One thread per entry slot; slot q belongs to atom q >> cap_log2.
__global__ __launch_bounds__(256) void k_nl_rev(Geo g, Work w)
{
    const long q = (long)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (q >= (long)w.nat * w.cap) return;
    const int i = (int)(q >> w.cap_log2);
    if ((int)(q - ((long)i << w.cap_log2)) >= w.count[i]) return;
    const StructGeo &G = g.sg[w.sid[i]];
    const int S = G.S;
    const int j = w.ent_j[q];
    const int sp = S - 1 - w.ent_s[q];
    const int bj = j * w.cap, nj = w.count[j];
    int r = -1;
    if (w.pos_start) {
        for (int t = w.pos_start[G.ps_off + (long long)(j - G.atom_off) * S + sp]; t < nj; ++t) {
            if (w.ent_s[bj + t] != sp) break;
            if (w.ent_j[bj + t] == i) {
                r = bj + t;
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

void launch_nl_rev(const Geo &g, Work &w, hipStream_t s)
{
    const long slots = (long)w.nat * w.cap;
    hipLaunchKernelGGL(k_nl_rev, dim3((unsigned)((slots + 255) / 256)), dim3(256), 0, s, g, w);
}

// Element segments of the element-sorted atom order: elem_off[e] = first
// position whose element is >= e (binary search), elem_count = differences.
__global__ void k_elem_bounds(Work w, const int *sorted)
{
    __shared__ int off[kMaxEl + 1];
    const int e = threadIdx.x;
    if (e <= kMaxEl) {
        int lo = 0, hi = w.nat;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (sorted[mid] < e) lo = mid + 1;
            else hi = mid;
        }
        off[e] = lo;
        w.elem_off[e] = lo;
    }
    __syncthreads();
    if (e < kMaxEl) w.elem_count[e] = off[e + 1] - off[e];
}

void launch_elem_bounds(Work &w, const int *sorted_elem, hipStream_t s)
{
    hipLaunchKernelGGL(k_elem_bounds, dim3(1), dim3(64), 0, s, w, sorted_elem);
}

}  // namespace hdnnp
