// Velocity-Verlet kernels of the on-device molecular dynamics and the Verlet
// skin bookkeeping.
//
// Reference: the MD loop of hdnnpExample.f90:58-72,
//   ats = ats + p / m * dt + 0.5 * forces / m * dt**2
//   lastForces = forces;  forces = F(ats)
//   p = p + dt / 2 * (forces + lastForces)
//   ekin = sum(p**2 / m) / 2
// with the same operation order per component.
#include "device.hpp"

namespace hdnnp {

// every product and sum rounded as written, so that the fused kernels of a
// reused-list step and the separate ones of a careful step agree bit for bit
#pragma clang fp contract(off)

// drift, keeping the forces of the step before in Fold for the kick
__global__ __launch_bounds__(256) void k_md_drift(double *pos, const double *mom, const double *F, double *Fold,
                                                  const double *mass, double dt, int n3)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= n3) return;
    const double m = mass[x / 3];
    const double f = F[x];
    pos[x] = (pos[x] + mom[x] / m * dt) + 0.5 * f / m * (dt * dt);
    Fold[x] = f;
}

__global__ __launch_bounds__(256) void k_md_kick(double *mom, const double *F, const double *Fold, double dt, int n3)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= n3) return;
    mom[x] = mom[x] + dt / 2.0 * (F[x] + Fold[x]);
}

// largest squared displacement since the last list build, into maxd2[*step]
// (maxd2[0] without a step counter); maxd2 holds the bits of non-negative
// doubles, which order like unsigned integers
__global__ __launch_bounds__(256) void k_md_disp(const double *pos, const double *posb, int nat,
                                                 unsigned long long *maxd2, const int *step)
{
    __shared__ double red[256];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double d2 = 0.0;
    if (i < nat) {
        const double a = pos[3 * i] - posb[3 * i], b = pos[3 * i + 1] - posb[3 * i + 1],
                     c = pos[3 * i + 2] - posb[3 * i + 2];
        d2 = a * a + b * b + c * c;
    }
    red[threadIdx.x] = d2;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(maxd2 + (step ? *step : 0), (unsigned long long)__double_as_longlong(red[0]));
}

// wrapped positions for a reused list: those of the build plus the
// displacement since (no re-wrapping, so the image shifts stay valid)
__global__ __launch_bounds__(256) void k_md_refresh(double4 *wpos, const double4 *wposb, const double *pos,
                                                    const double *posb, int nat)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nat) return;
    const double4 b = wposb[i];
    wpos[i] = make_double4(b.x + (pos[3 * i] - posb[3 * i]), b.y + (pos[3 * i + 1] - posb[3 * i + 1]),
                           b.z + (pos[3 * i + 2] - posb[3 * i + 2]), 0.0);
}

// The start of a step on reused lists, one thread per atom: k_md_drift,
// k_md_disp (into maxd2[*step]) and k_md_refresh in one pass, with their
// arithmetic.
__global__ __launch_bounds__(256) void k_md_begin(double *pos, const double *mom, const double *F, double *Fold,
                                                  const double *mass, double dt, const double *posb, double4 *wpos,
                                                  const double4 *wposb, int nat, unsigned long long *maxd2,
                                                  const int *step)
{
    __shared__ double red[256];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double d2 = 0.0;
    if (i < nat) {
        const double m = mass[i];
        double p[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const int x = 3 * i + c;
            const double f = F[x];
            p[c] = (pos[x] + mom[x] / m * dt) + 0.5 * f / m * (dt * dt);
            pos[x] = p[c];
            Fold[x] = f;
        }
        const double a = p[0] - posb[3 * i], b = p[1] - posb[3 * i + 1], c = p[2] - posb[3 * i + 2];
        d2 = a * a + b * b + c * c;
        const double4 w = wposb[i];
        wpos[i] = make_double4(w.x + a, w.y + b, w.z + c, 0.0);
    }
    red[threadIdx.x] = d2;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(maxd2 + *step, (unsigned long long)__double_as_longlong(red[0]));
}

// The end of a step: k_md_kick, then k_md_save of the state after step *step.
__global__ __launch_bounds__(256) void k_md_end(double *mom, const double *F, const double *Fold, double dt,
                                                double *slot, const double *pos, int n3, const int *step)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= n3) return;
    const double p = mom[x] + dt / 2.0 * (F[x] + Fold[x]);
    mom[x] = p;
    slot += (size_t)(*step + 1) * 3 * n3;
    slot[x] = pos[x];
    slot[n3 + x] = p;
    slot[2 * n3 + x] = F[x];
}

// state of a step (positions, momenta, forces) to or from one saved slot;
// with a step counter the slot after step *step (slots of 3 n3 doubles)
__global__ __launch_bounds__(256) void k_md_save(double *slot, const double *pos, const double *mom, const double *F,
                                                 int n3, const int *step)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= n3) return;
    if (step) slot += (size_t)(*step + 1) * 3 * n3;
    slot[x] = pos[x];
    slot[n3 + x] = mom[x];
    slot[2 * n3 + x] = F[x];
}

__global__ __launch_bounds__(256) void k_md_load(const double *slot, double *pos, double *mom, double *F, int n3)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= n3) return;
    pos[x] = slot[x];
    mom[x] = slot[n3 + x];
    F[x] = slot[2 * n3 + x];
}

// the step counter of a chunk, advanced at the end of each step
__global__ void k_md_tick(int *step)
{
    if (threadIdx.x == 0) *step += 1;
}

// kinetic energy of each system: one block per system, fixed-order tree
__global__ __launch_bounds__(256) void k_md_ekin(const double *mom, const double *mass, const int *off,
                                                 double *ekin)
{
    __shared__ double red[256];
    const int s = blockIdx.x;
    double a = 0.0;
    for (int i = off[s] + threadIdx.x; i < off[s + 1]; i += 256) {
        const double m = mass[i];
        a += mom[3 * i] * mom[3 * i] / m + mom[3 * i + 1] * mom[3 * i + 1] / m +
             mom[3 * i + 2] * mom[3 * i + 2] / m;
    }
    red[threadIdx.x] = a;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) ekin[s] = red[0] / 2.0;
}

void launch_md_drift(double *pos, const double *mom, const double *F, double *Fold, const double *mass, double dt,
                     int nat, hipStream_t s)
{
    if (nat > 0)
        hipLaunchKernelGGL(k_md_drift, dim3((3 * nat + 255) / 256), dim3(256), 0, s, pos, mom, F, Fold, mass, dt, 3 * nat);
}

void launch_md_kick(double *mom, const double *F, const double *Fold, double dt, int nat, hipStream_t s)
{
    if (nat > 0) hipLaunchKernelGGL(k_md_kick, dim3((3 * nat + 255) / 256), dim3(256), 0, s, mom, F, Fold, dt, 3 * nat);
}

void launch_md_disp(const double *pos, const double *posb, int nat, unsigned long long *maxd2, const int *step,
                    hipStream_t s)
{
    if (nat > 0) hipLaunchKernelGGL(k_md_disp, dim3((nat + 255) / 256), dim3(256), 0, s, pos, posb, nat, maxd2, step);
}

void launch_md_refresh(double4 *wpos, const double4 *wposb, const double *pos, const double *posb, int nat,
                       hipStream_t s)
{
    if (nat > 0) hipLaunchKernelGGL(k_md_refresh, dim3((nat + 255) / 256), dim3(256), 0, s, wpos, wposb, pos, posb, nat);
}

void launch_md_save(double *slot, const double *pos, const double *mom, const double *F, int nat, const int *step,
                    hipStream_t s)
{
    if (nat > 0)
        hipLaunchKernelGGL(k_md_save, dim3((3 * nat + 255) / 256), dim3(256), 0, s, slot, pos, mom, F, 3 * nat, step);
}

void launch_md_begin(double *pos, const double *mom, const double *F, double *Fold, const double *mass, double dt,
                     const double *posb, double4 *wpos, const double4 *wposb, int nat, unsigned long long *maxd2,
                     const int *step, hipStream_t s)
{
    if (nat > 0)
        hipLaunchKernelGGL(k_md_begin, dim3((nat + 255) / 256), dim3(256), 0, s, pos, mom, F, Fold, mass, dt, posb, wpos,
                           wposb, nat, maxd2, step);
}

void launch_md_end(double *mom, const double *F, const double *Fold, double dt, double *slot, const double *pos,
                   int nat, const int *step, hipStream_t s)
{
    if (nat > 0)
        hipLaunchKernelGGL(k_md_end, dim3((3 * nat + 255) / 256), dim3(256), 0, s, mom, F, Fold, dt, slot, pos, 3 * nat,
                           step);
}

void launch_md_tick(int *step, hipStream_t s) { hipLaunchKernelGGL(k_md_tick, dim3(1), dim3(64), 0, s, step); }

void launch_md_load(const double *slot, double *pos, double *mom, double *F, int nat, hipStream_t s)
{
    if (nat > 0) hipLaunchKernelGGL(k_md_load, dim3((3 * nat + 255) / 256), dim3(256), 0, s, slot, pos, mom, F, 3 * nat);
}

void launch_md_ekin(const double *mom, const double *mass, const int *off, int nsys, double *ekin, hipStream_t s)
{
    if (nsys > 0) hipLaunchKernelGGL(k_md_ekin, dim3(nsys), dim3(256), 0, s, mom, mass, off, ekin);
}

}  // namespace hdnnp
