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
// side only): add their contribution to the neighbour directly. One thread
// per entry slot; slot q belongs to atom q >> cap_log2.
template <typename R>
__global__ __launch_bounds__(256) void k_orphans(Work w, const R *g)
{
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (long)w.nat * w.cap) return;
    const int i = (int)(q >> w.cap_log2);
    if ((int)(q - ((long)i << w.cap_log2)) >= w.count[i] || w.rev[q] >= 0) return;
    const int j = w.ent_j[q];
    atomicAdd(&w.force[3 * j + 0], -(double)g[3 * (size_t)q + 0]);
    atomicAdd(&w.force[3 * j + 1], -(double)g[3 * (size_t)q + 1]);
    atomicAdd(&w.force[3 * j + 2], -(double)g[3 * (size_t)q + 2]);
}

template <typename R>
void launch_force_gather_t(Work &w, const R *gent, hipStream_t s)
{
    hipLaunchKernelGGL(k_force_gather<R>, dim3(w.nat), dim3(64), 0, s, w, gent);
    const long slots = (long)w.nat * w.cap;
    hipLaunchKernelGGL(k_orphans<R>, dim3((unsigned)((slots + 255) / 256)), dim3(256), 0, s, w, gent);
}
template void launch_force_gather_t<double>(Work &, const double *, hipStream_t);
template void launch_force_gather_t<float>(Work &, const float *, hipStream_t);

// Two-stage FP64 reduction, result[0] = sum Ei, result[1..9] = sum vir.
// Fixed order for a given nat: block b sums a fixed contiguous chunk (thread
// t takes every 256th atom of it, then a tree over threads); one block then
// sums the block partials with the same tree.
constexpr int kRedBlocks = 256;

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

__global__ __launch_bounds__(256) void k_reduce_partial(Work w, double *partial, int chunk)
{
    __shared__ double red[10][256];
    double a[10];
#pragma unroll
    for (int k = 0; k < 10; ++k) a[k] = 0;
    const int i0 = blockIdx.x * chunk, i1 = min(w.nat, i0 + chunk);
    for (int i = i0 + threadIdx.x; i < i1; i += 256) {
        a[0] += w.Ei[i];
#pragma unroll
        for (int k = 0; k < 9; ++k) a[1 + k] += w.vir[9 * (size_t)i + k];
    }
    tree10(red, a);
    if (threadIdx.x < 10) partial[10 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

__global__ __launch_bounds__(256) void k_reduce_final(Work w, const double *partial, int nb)
{
    __shared__ double red[10][256];
    double a[10];
#pragma unroll
    for (int k = 0; k < 10; ++k) a[k] = (int)threadIdx.x < nb ? partial[10 * threadIdx.x + k] : 0.0;
    tree10(red, a);
    if (threadIdx.x < 10) w.result[threadIdx.x] = red[threadIdx.x][0];
}

void launch_reduce(Work &w, double *partial, hipStream_t s)
{
    const int nb = std::min(kRedBlocks, (w.nat + 255) / 256);
    const int chunk = (w.nat + nb - 1) / nb;
    hipLaunchKernelGGL(k_reduce_partial, dim3(nb), dim3(256), 0, s, w, partial, chunk);
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(256), 0, s, w, (const double *)partial, nb);
}

}  // namespace hdnnp
