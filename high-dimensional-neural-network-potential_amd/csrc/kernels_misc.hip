// Force gather and the FP64 energy / virial reductions.
//
// Reference: the force scatter of hdnnpCalcNNs
// (src/potentials/hdnnps.f90:533-538) and E = sum(atEns) (:557). Instead of
// scattering with atomics, each atom gathers: its own entries' gradient rows
// (the self slot, whose gradient is minus the sum of the others) and the
// partner rows of the neighbours that see it, in a fixed order, so results are
// reproducible bit for bit.
#include "device.hpp"

#include <algorithm>

namespace hdnnp {

// One wave per atom m, lanes over its entries: F_m = sum_q (g[q] - g[rev[q]]),
// where g[q] = dE_m/dR_j of entry q and g[rev[q]] = dE_j/dR_m of its partner.
// The lane sums are combined by a fixed butterfly, so the result does not
// depend on scheduling.
template <typename R>
__global__ __launch_bounds__(64) void k_force_gather(Work w, const R *g)
{
    const int m = xcd_block(blockIdx.x, gridDim.x);
    const int lane = threadIdx.x;
    const int o = m * w.cap, n = w.count[m];
    double fx = 0, fy = 0, fz = 0;
    for (int q = o + lane; q < o + n; q += 64) {
        fx += (double)g[3 * (size_t)q + 0];
        fy += (double)g[3 * (size_t)q + 1];
        fz += (double)g[3 * (size_t)q + 2];
        const int r = w.rev[q];
        if (r >= 0) {
            fx -= (double)g[3 * (size_t)r + 0];
            fy -= (double)g[3 * (size_t)r + 1];
            fz -= (double)g[3 * (size_t)r + 2];
        }
    }
#pragma unroll
    for (int k = 32; k > 0; k >>= 1) {
        fx += __shfl_xor(fx, k);
        fy += __shfl_xor(fy, k);
        fz += __shfl_xor(fz, k);
    }
    if (lane == 0) {
        w.force[3 * m + 0] = fx;
        w.force[3 * m + 1] = fy;
        w.force[3 * m + 2] = fz;
    }
}

// entries whose partner is missing (pair at the exact cutoff seen from one
// side only): add their contribution to the neighbour directly. One
// wavefront per atom (four per block), its lanes over the atom's entries.
template <typename R>
__device__ void orphans_block(const Work &w, const R *g, int blk)
{
    const int i = blk * 4 + (int)threadIdx.x / 64;
    if (i >= w.nat) return;
    const int cnt = w.count[i];
    for (int t = (int)threadIdx.x % 64; t < cnt; t += 64) {
        const size_t q = (size_t)i * w.cap + t;
        if (w.rev[q] >= 0) continue;
        const int j = w.ent_j[q];
        atomicAdd(&w.force[3 * j + 0], -(double)g[3 * q + 0]);
        atomicAdd(&w.force[3 * j + 1], -(double)g[3 * q + 1]);
        atomicAdd(&w.force[3 * j + 2], -(double)g[3 * q + 2]);
    }
}

// the orphans run in the launch of the first reduction stage (launch_reduce)
template <typename R>
void launch_force_gather_t(Work &w, const R *gent, hipStream_t s)
{
    hipLaunchKernelGGL(k_force_gather<R>, dim3(w.nat), dim3(64), 0, s, w, gent);
}
template void launch_force_gather_t<double>(Work &, const double *, hipStream_t);
template void launch_force_gather_t<float>(Work &, const float *, hipStream_t);

// Two-stage FP64 reduction per structure: result[10 s] = sum of Ei,
// result[10 s + 1 + k] = sum of vir[k] over the atoms of structure s.
// Stage 1: one block per chunk of atoms (chunks never cross a structure;
// thread t takes every 256th atom of the chunk, then a tree over threads).
// Stage 2: one wave per structure adds its chunks' partials in a fixed order.
// The order depends only on the structure's atom count, so results are
// reproducible and a structure gets the same sums alone or in a batch.
__device__ inline void tree10(double (*red)[256], double *a)
{
#pragma unroll
    for (int k = 0; k < 10; ++k) red[k][threadIdx.x] = a[k];
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o)
#pragma unroll
            for (int k = 0; k < 10; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + o];
        __syncthreads();
    }
}

__device__ void reduce_final_wave(const Work &w, const int *sc, const double *partial, int *flag_out, int s,
                                  int lane);

// FUSE (few structures): the block that finishes last, by an atomic count of
// finished blocks in `done`, runs stage 2 itself (the same per-wave code in
// the same order, its four waves over the structures) and sets the count back
// to zero, so that one launch does both stages
template <bool FUSE>
__device__ void reduce_partial_block(const Work &w, const int2 *chunks, double *partial, const int *sc, int nstruct,
                                     int *flag_out, unsigned *done, int blk, int nchunk)
{
    __shared__ double red[10][256];
    double a[10];
#pragma unroll
    for (int k = 0; k < 10; ++k) a[k] = 0;
    const int2 c = chunks[blk];
    for (int i = c.x + threadIdx.x; i < c.y; i += 256) {
        a[0] += w.Ei[i];
#pragma unroll
        for (int k = 0; k < 9; ++k) a[1 + k] += w.vir[9 * (size_t)i + k];
    }
    tree10(red, a);
    if (threadIdx.x < 10) partial[10 * (size_t)blk + threadIdx.x] = red[threadIdx.x][0];
    if constexpr (FUSE) {
        __shared__ int last;
        __threadfence();   // this block's partials reach L2 before it is counted
        __syncthreads();
        if (threadIdx.x == 0) last = atomicAdd(done, 1u) == (unsigned)nchunk - 1;
        __syncthreads();
        if (!last) return;
        __threadfence();   // and the other blocks' are read from there
        for (int st = (int)threadIdx.x / 64; st < nstruct; st += 4)
            reduce_final_wave(w, sc, partial, flag_out, st, (int)threadIdx.x % 64);
        if (threadIdx.x == 0) *done = 0;
    }
}

template <bool FUSE>
__global__ __launch_bounds__(256) void k_reduce_partial(Work w, const int2 *chunks, double *partial, const int *sc,
                                                        int nstruct, int *flag_out, unsigned *done)
{
    reduce_partial_block<FUSE>(w, chunks, partial, sc, nstruct, flag_out, done, (int)blockIdx.x, (int)gridDim.x);
}

// blocks [0, norph): the orphans of four atoms each; the blocks after: the
// first reduction stage, one chunk each (independent of the forces, so the
// two share a launch)
template <typename R, bool FUSE>
__global__ __launch_bounds__(256) void k_orphans_reduce(Work w, const R *g, int norph, const int2 *chunks,
                                                        double *partial, const int *sc, int nstruct, int *flag_out,
                                                        unsigned *done, int nchunk)
{
    if ((int)blockIdx.x < norph) {   // block-uniform
        orphans_block<R>(w, g, (int)blockIdx.x);
        return;
    }
    reduce_partial_block<FUSE>(w, chunks, partial, sc, nstruct, flag_out, done, (int)blockIdx.x - norph, nchunk);
}

// one wave per structure: lane l adds chunks l, l+64, ... of the structure,
// then a fixed butterfly over the lanes
// flag_out (block 0, when given) receives a copy of the 8 flag words, so that
// the host reads results and flags back in one copy
__device__ void reduce_final_wave(const Work &w, const int *sc, const double *partial, int *flag_out, int s,
                                  int lane)
{
    double a[10];
#pragma unroll
    for (int k = 0; k < 10; ++k) a[k] = 0;
    for (int c = sc[s] + lane; c < sc[s + 1]; c += 64)
#pragma unroll
        for (int k = 0; k < 10; ++k) a[k] += partial[10 * (size_t)c + k];
#pragma unroll
    for (int k = 0; k < 10; ++k)
        for (int o = 32; o > 0; o >>= 1) a[k] += __shfl_xor(a[k], o);
    if (lane < 10) {
        double v = a[0];
#pragma unroll
        for (int k = 1; k < 10; ++k) v = lane == k ? a[k] : v;
        w.result[10 * (size_t)s + lane] = v;
    }
    if (flag_out && s == 0 && lane >= 16 && lane < 24) flag_out[lane - 16] = w.flags[lane - 16];
}

__global__ __launch_bounds__(64) void k_reduce_final(Work w, const int *sc, const double *partial, int *flag_out)
{
    reduce_final_wave(w, sc, partial, flag_out, (int)blockIdx.x, (int)threadIdx.x);
}

// up to this many structures and chunks, one launch does both stages (FUSE):
// each block's device-scope fence writes its XCD's L2 back, which costs more
// than the second launch once there are many blocks (96 chunks: 14.2 against
// 9.3 us for the two launches)
constexpr int kFuseReduceMax = 16;

template <typename R, bool FUSE>
static void launch_stage1(Work &w, const R *og, const int2 *chunks, int nchunk, const int *sc, int nstruct,
                          double *partial, int *flag_out, unsigned *done, hipStream_t s)
{
    if (og) {
        const int norph = (w.nat + 3) / 4;
        hipLaunchKernelGGL((k_orphans_reduce<R, FUSE>), dim3(norph + nchunk), dim3(256), 0, s, w, og, norph, chunks,
                           partial, sc, nstruct, flag_out, done, nchunk);
    } else if (nchunk > 0) {
        hipLaunchKernelGGL(k_reduce_partial<FUSE>, dim3(nchunk), dim3(256), 0, s, w, chunks, partial, sc, nstruct,
                           flag_out, done);
    }
}

template <typename R>
void launch_reduce(Work &w, const int2 *chunks, int nchunk, const int *sc, int nstruct, double *partial,
                   int *flag_out, unsigned *done, const R *orphans_g, hipStream_t s)
{
    if (nchunk > 0 && nstruct <= kFuseReduceMax && nchunk <= kFuseReduceMax) {
        launch_stage1<R, true>(w, orphans_g, chunks, nchunk, sc, nstruct, partial, flag_out, done, s);
        return;
    }
    launch_stage1<R, false>(w, orphans_g, chunks, nchunk, sc, nstruct, partial, flag_out, done, s);
    hipLaunchKernelGGL(k_reduce_final, dim3(nstruct), dim3(64), 0, s, w, sc, (const double *)partial, flag_out);
}
template void launch_reduce<double>(Work &, const int2 *, int, const int *, int, double *, int *, unsigned *,
                                    const double *, hipStream_t);
template void launch_reduce<float>(Work &, const int2 *, int, const int *, int, double *, int *, unsigned *,
                                   const float *, hipStream_t);

}  // namespace hdnnp
