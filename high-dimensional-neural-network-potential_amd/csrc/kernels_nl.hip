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

// Element rank of each atom; unknown Z flagged (rare, so the atomic is cheap).
__global__ __launch_bounds__(256) void k_map_elements(DevModel m, Work w)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= w.nat) return;
    const int z = w.Z[i];
    const int e = (z >= 0 && z < 128) ? m.zlut[z] : -1;
    if (e < 0) atomicMax(&w.flags[0], z > 0 ? z : 1);
    w.eidx[i] = e < 0 ? 0 : e;
}

void launch_map_elements(const DevModel &m, Work &w, hipStream_t s)
{
    hipLaunchKernelGGL(k_map_elements, dim3((w.nat + 255) / 256), dim3(256), 0, s, m, w);
}

// Bounding box of the raw positions (nonperiodic artificial lattice,
// neighborLists.f90:131-138): bbox[0..2] = min, bbox[3..5] = max.
__global__ __launch_bounds__(1024) void k_bbox(Work w)
{
    __shared__ double red[6][1024];
    double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
    for (int i = threadIdx.x; i < w.nat; i += blockDim.x)
        for (int k = 0; k < 3; ++k) {
            const double v = w.pos[3 * i + k];
            mn[k] = fmin(mn[k], v);
            mx[k] = fmax(mx[k], v);
        }
    for (int k = 0; k < 3; ++k) {
        red[k][threadIdx.x] = mn[k];
        red[3 + k][threadIdx.x] = mx[k];
    }
    __syncthreads();
    for (int o = 512; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o)
            for (int k = 0; k < 3; ++k) {
                red[k][threadIdx.x] = fmin(red[k][threadIdx.x], red[k][threadIdx.x + o]);
                red[3 + k][threadIdx.x] = fmax(red[3 + k][threadIdx.x], red[3 + k][threadIdx.x + o]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) w.bbox[threadIdx.x] = red[threadIdx.x][0];
}

void launch_bbox(Work &w, hipStream_t s)
{
    hipLaunchKernelGGL(k_bbox, dim3(1), dim3(1024), 0, s, w);
}

// Wrap into the list frame and bin: periodic input is wrapped twice (as
// hdnnps.f90:381 and neighborLists.f90:156 do); nonperiodic input is shifted
// by the artificial origin and wrapped once in the artificial box. Cell index
// int(s*ngrid) (neighborLists.f90:161-166), clamped into the grid.
__global__ __launch_bounds__(256) void k_wrap_bin(Geo g, Work w)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= w.nat) return;
    double x = w.pos[3 * i + 0], y = w.pos[3 * i + 1], z = w.pos[3 * i + 2];
    if (g.periodic) {
        wrap1(g.lat, g.inv, x, y, z);
        wrap1(g.lat, g.inv, x, y, z);
    } else {
        x = x - g.origin[0];
        y = y - g.origin[1];
        z = z - g.origin[2];
        wrap1(g.lat, g.inv, x, y, z);
    }
    w.wpos[i] = make_double4(x, y, z, 0.0);
    double s0, s1, s2;
    mv3(g.inv, x, y, z, s0, s1, s2);
    int c0 = (int)(s0 * g.ng[0]), c1 = (int)(s1 * g.ng[1]), c2 = (int)(s2 * g.ng[2]);
    c0 = min(max(c0, 0), g.ng[0] - 1);
    c1 = min(max(c1, 0), g.ng[1] - 1);
    c2 = min(max(c2, 0), g.ng[2] - 1);
    w.cellkey[i] = (c0 * g.ng[1] + c1) * g.ng[2] + c2;
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

// One stencil cell of atom i: wrapped cell id and image shift, in the
// reference's order of wrapping (z, then y, then x; neighborLists.f90:186-209).
// Returns false when the cell lies outside a nonperiodic grid.
struct StencilCell {
    int cell;
    int sx, sy, sz;
    double dl0, dl1, dl2;
    bool centre;
};

__device__ inline bool stencil_cell(const Geo &g, int key, int s, StencilCell &c)
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

// the strict distance test of neighborLists.f90:213-214
__device__ inline bool is_neighbor(const Geo &g, const double4 &pi, const double4 &pj,
                                   const StencilCell &c)
{
    const double d0 = pi.x - (pj.x + c.dl0);
    const double d1 = pi.y - (pj.y + c.dl1);
    const double d2 = pi.z - (pj.z + c.dl2);
    return (d0 * d0 + d1 * d1) + d2 * d2 < g.rc2;
}

__device__ inline int wave_sum_int(int v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__device__ inline int wave_excl_scan_int(int v, int lane)
{
    int x = v;
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    return x - v;
}

// Count pass: one wave per atom, lanes over the stencil cells.
__global__ __launch_bounds__(256) void k_nl_count(Geo g, Work w)
{
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= w.nat) return;
    const double4 pi = w.wpos[i];
    const int key = w.cellkey[i];
    const int S = (2 * g.nc[0] + 1) * (2 * g.nc[1] + 1) * (2 * g.nc[2] + 1);
    int cnt = 0;
    for (int s = lane; s < S; s += 64) {
        StencilCell c;
        if (!stencil_cell(g, key, s, c)) continue;
        for (int q = w.cell_start[c.cell]; q < w.cell_end[c.cell]; ++q) {
            const int j = w.cell_atoms[q];
            if (j == i && c.centre) continue;
            cnt += is_neighbor(g, pi, w.wpos[j], c) ? 1 : 0;
        }
    }
    cnt = wave_sum_int(cnt);
    if (lane == 0) w.count[i] = cnt;
}

void launch_nl_count(const Geo &g, Work &w, hipStream_t s)
{
    hipLaunchKernelGGL(k_nl_count, dim3((w.nat + 3) / 4), dim3(256), 0, s, g, w);
}

// Fill pass: the same enumeration; entries are written in stencil order and,
// within a cell, in atom order -- the reference's list order (self excluded).
__global__ __launch_bounds__(256) void k_nl_fill(Geo g, Work w)
{
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= w.nat) return;
    const double4 pi = w.wpos[i];
    const int key = w.cellkey[i];
    const int S = (2 * g.nc[0] + 1) * (2 * g.nc[1] + 1) * (2 * g.nc[2] + 1);
    int base = w.offset[i];
    for (int s0 = 0; s0 < S; s0 += 64) {
        const int s = s0 + lane;
        StencilCell c;
        const bool ok = s < S && stencil_cell(g, key, s, c);
        int cnt = 0;
        if (ok)
            for (int q = w.cell_start[c.cell]; q < w.cell_end[c.cell]; ++q) {
                const int j = w.cell_atoms[q];
                if (j == i && c.centre) continue;
                cnt += is_neighbor(g, pi, w.wpos[j], c) ? 1 : 0;
            }
        int pos = base + wave_excl_scan_int(cnt, lane);
        if (ok && cnt) {
            if (c.sx < -127 || c.sx > 127 || c.sy < -127 || c.sy > 127 || c.sz < -127 || c.sz > 127)
                atomicMax(&w.flags[2], 1);
            const int sh = pack_shift(c.sx, c.sy, c.sz);
            for (int q = w.cell_start[c.cell]; q < w.cell_end[c.cell]; ++q) {
                const int j = w.cell_atoms[q];
                if (j == i && c.centre) continue;
                if (is_neighbor(g, pi, w.wpos[j], c)) {
                    w.ent_j[pos] = j;
                    w.ent_shift[pos] = sh;
                    ++pos;
                }
            }
        }
        base += wave_sum_int(cnt);
    }
}

void launch_nl_fill(const Geo &g, Work &w, hipStream_t s)
{
    hipLaunchKernelGGL(k_nl_fill, dim3((w.nat + 3) / 4), dim3(256), 0, s, g, w);
}

// Partner entries: entry (i -> j, L) pairs with (j -> i, -L) in j's list.
// The strict test is evaluated separately in each direction, so a pair at
// the exact cutoff can lack its partner; such entries get rev = -1.
__global__ __launch_bounds__(256) void k_nl_rev(Work w)
{
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= w.nat) return;
    const int o = w.offset[i], n = w.count[i];
    for (int q = lane; q < n; q += 64) {
        const int j = w.ent_j[o + q];
        int x, y, z;
        unpack_shift(w.ent_shift[o + q], x, y, z);
        const int want = pack_shift(-x, -y, -z);
        const int oj = w.offset[j], nj = w.count[j];
        int r = -1;
        for (int t = 0; t < nj; ++t)
            if (w.ent_j[oj + t] == i && w.ent_shift[oj + t] == want) {
                r = oj + t;
                break;
            }
        w.rev[o + q] = r;
    }
}

void launch_nl_rev(Work &w, hipStream_t s)
{
    hipLaunchKernelGGL(k_nl_rev, dim3((w.nat + 3) / 4), dim3(256), 0, s, w);
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
