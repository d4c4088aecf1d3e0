// Per-element MLP on the matrix cores: forward pass and backpropagation of
// dE/dG for tiles of 16 atoms of one element.
//
// Reference: nn_forward / layer_forward and nn_backward / layer_backward
// (src/neuralnetwork/neuralNetworks.f90:94-158, 220-313) with activations
// activate / deactivate (:33-84), called per atom by hdnnpCalcNNs
// (src/potentials/hdnnps.f90:523-531) with gradOut = 1.
//
// One wavefront per tile. A layer is C[16 x nout] = X[16 x nin] . W[nin x nout]
// computed with v_mfma_{f32,f64}_16x16x4 (FP32 / FP64 accumulation), 16 output
// columns per accumulator; the backward pass is G[16 x nin] = D[16 x nout] .
// W^T. Operand layout of the 16x16x4 instructions (wave64): lane l supplies
// A[l%16][l/16] and B[l/16][l%16]; it holds D[4*(l/16)+r][l%16] (f32) or
// D[(l/16)+4r][l%16] (f64), r = 0..3 (checked on gfx950 with an in-bounds
// probe of both instructions).
// The tile's activations and pre-activations stay in LDS (row stride odd, so
// the 16 rows of one column fall in distinct banks).
#include "device.hpp"
#include "fmath.hpp"

namespace hdnnp {

template <typename R> struct Mma;
template <> struct Mma<float> {
    typedef float V __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ V op(float a, float b, V c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int kq, int r) { return 4 * kq + r; }
};
template <> struct Mma<double> {
    typedef double V __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ V op(double a, double b, V c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int kq, int r) { return kq + 4 * r; }
};

__device__ __forceinline__ double act_f(char a, double x)
{
    switch (a) {
    case 't': return tanh_f64(x);
    case 'p': return log(1.0 + exp(x));
    case 'r': return x > 0.0 ? x : 0.0;
    default: return x;
    }
}
__device__ __forceinline__ double act_d(char a, double x)
{
    switch (a) {
    case 't': { const double t = tanh_f64(x); return 1.0 - t * t; }
    case 'p': return 1.0 / (1.0 + exp(-x));
    case 'r': return x > 0.0 ? 1.0 : 0.0;
    default: return 1.0;
    }
}
__device__ __forceinline__ float act_f(char a, float x)
{
    switch (a) {
    case 't': return tanhf(x);
    case 'p': return logf(1.0f + __expf(x));
    case 'r': return x > 0.0f ? x : 0.0f;
    default: return x;
    }
}
__device__ __forceinline__ float act_d(char a, float x)
{
    switch (a) {
    case 't': { const float t = tanhf(x); return 1.0f - t * t; }
    case 'p': return 1.0f / (1.0f + __expf(-x));
    case 'r': return x > 0.0f ? 1.0f : 0.0f;
    default: return 1.0f;
    }
}

// BF16 operands for the forward layers (v_mfma_f32_16x16x16_bf16, FP32
// accumulation): lane (ri, kq) supplies X[ri][k0 + 4 kq + i] and
// W[k0 + 4 kq + i][j0 + ri], i = 0..3, and holds D[4 kq + r][j0 + ri] as the
// FP32 16x16x4 form does. Rounding to nearest even.
typedef short Bf4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ short bf16_rne(float x)
{
    const unsigned u = __float_as_uint(x);
    return (short)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

template <typename R> __device__ __forceinline__ const R *weights(const DevModel &m, int e);
template <> __device__ __forceinline__ const double *weights<double>(const DevModel &m, int e) { return m.W64[e]; }
template <> __device__ __forceinline__ const float *weights<float>(const DevModel &m, int e) { return m.W32[e]; }

// Layer geometry passed by value: nodes, activations and LDS offsets of the
// pre-activations (only indexed with loop counters the compiler unrolls).
struct MlpShape {
    int nlayers;
    int nodes[kMaxLay];
    char act[kMaxLay];
    int ldc[kMaxLay];       // row stride of the pre-activations of layer l
    int coff[kMaxLay];      // their offset (elements) in LDS
    int ld;                 // row stride of the two activation buffers
    int total;              // LDS elements
};

constexpr int kKU = 8;   // matrix-core steps whose operands are loaded together
// output column groups of 16 computed together (their loads go out at once)
template <typename R> constexpr int kJG = sizeof(R) == 4 ? 2 : 1;

// Ei_out[i] = E_i of this network; dEdG = scale * dE_i/dG, or added to dEdG
// when accumulate is set (committee mean over several networks). BF16 (FP32
// only): the forward layers take BF16 operands; the backward pass stays FP32.
template <typename R, bool BF16>
__global__ __launch_bounds__(64) void k_mlp(DevModel m, Work w, MlpShape sh, const R *G, R *dEdG, double *Ei_out,
                                            R scale, int accumulate)
{
    extern __shared__ __align__(16) char smem[];
    typedef typename Mma<R>::V V;
    constexpr int T = 16;
    int e = -1, t0 = 0, tn = 0;
    {
        int acc = 0;
        for (int f = 0; f < m.nel; ++f) {
            const int cnt = w.elem_count[f];
            const int nt = (cnt + T - 1) / T;
            if ((int)blockIdx.x < acc + nt) {
                const int t = blockIdx.x - acc;
                e = f;
                t0 = w.elem_off[f] + t * T;
                tn = min(T, cnt - t * T);
                break;
            }
            acc += nt;
        }
    }
    if (e < 0) return;
    const int lane = threadIdx.x;
    const int ri = lane & 15, kq = lane >> 4;
    const int nsf = m.einfo[e].nsf;
    const int ldg = m.nsfmax;
    const int ld = sh.ld;
    R *P0 = (R *)smem;
    R *P1 = P0 + T * ld;
    R *Cb = P1 + T * ld;
    const R *W = weights<R>(m, e);
    // the atoms of the tile rows, in LDS past the activations: the loops
    // below read them there, so that their global loads do not wait on one
    // another
    int *rows = (int *)(smem + sizeof(R) * (size_t)sh.total);
    const int atom = lane < tn ? w.perm[t0 + lane] : 0;
    if (lane < T) rows[lane] = atom;
    __syncthreads();
    // lane (ri, kq) takes operand rows from the atom of tile row ri
    const R *Grow = G + (size_t)rows[ri] * ldg;
    const bool live = ri < tn;

    // forward, layer_forward (neuralNetworks.f90:94-114): c = W x + b, y = act(c);
    // the first layer reads the symmetry functions from global memory, the
    // others the activations of the layer before from LDS
    const R *Wl[kMaxLay];
    R *in = P1, *out = P0;
    int nin = nsf;
    const R *wp = W;
#pragma unroll
    for (int l = 0; l < kMaxLay; ++l) {
        if (l >= sh.nlayers) break;
        const int nout = sh.nodes[l];
        Wl[l] = wp;
        const R *bias = wp + (size_t)nin * nout;
        R *c = Cb + sh.coff[l];
        const int ldc = sh.ldc[l];
        const char act = sh.act[l];
        if constexpr (BF16 && sizeof(R) == 4) {
        for (int j0 = 0; j0 < nout; j0 += 16) {
            V acc = {0, 0, 0, 0};
            const int j = j0 + ri;
            {
                // 16 k per instruction, two instructions per step
                for (int k0 = 0; k0 < nin; k0 += 32) {
                    Bf4 a[2], b[2];
#pragma unroll
                    for (int u = 0; u < 2; ++u)
#pragma unroll
                        for (int i = 0; i < 4; ++i) {
                            const int k = k0 + 16 * u + 4 * kq + i;
                            const R x = k < nin ? (l == 0 ? (live ? Grow[k] : R(0)) : in[ri * ld + k]) : R(0);
                            const R y = (k < nin && j < nout) ? wp[(size_t)k * nout + j] : R(0);
                            a[u][i] = bf16_rne(x);
                            b[u][i] = bf16_rne(y);
                        }
#pragma unroll
                    for (int u = 0; u < 2; ++u) acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[u], b[u], acc, 0, 0, 0);
                }
            }
            if (j < nout) {
                const R bj = bias[j];
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = Mma<R>::row(kq, r);
                    const R cv = acc[r] + bj;
                    c[row * ldc + j] = cv;
                    out[row * ld + j] = act_f(act, cv);
                }
            }
        }
        } else {
        // kJG column groups of 16 at a time, kKU steps at a time: the
        // operands of a step are loaded once for all the groups, and the
        // loads of all (groups x steps) go out before their MFMAs; steps past
        // nin multiply zeros, which leaves the accumulators as they are
        for (int j00 = 0; j00 < nout; j00 += 16 * kJG<R>) {
            V acc[kJG<R>];
#pragma unroll
            for (int g = 0; g < kJG<R>; ++g) acc[g] = V{0, 0, 0, 0};
            for (int k0 = 0; k0 < nin; k0 += 4 * kKU) {
                R a[kKU], b[kJG<R>][kKU];
#pragma unroll
                for (int u = 0; u < kKU; ++u) {
                    const int k = k0 + 4 * u + kq;
                    a[u] = k < nin ? (l == 0 ? (live ? Grow[k] : R(0)) : in[ri * ld + k]) : R(0);
#pragma unroll
                    for (int g = 0; g < kJG<R>; ++g) {
                        const int j = j00 + 16 * g + ri;
                        b[g][u] = (k < nin && j < nout) ? wp[(size_t)k * nout + j] : R(0);
                    }
                }
#pragma unroll
                for (int g = 0; g < kJG<R>; ++g)
                    if (j00 + 16 * g < nout)
#pragma unroll
                        for (int u = 0; u < kKU; ++u) acc[g] = Mma<R>::op(a[u], b[g][u], acc[g]);
            }
#pragma unroll
            for (int g = 0; g < kJG<R>; ++g) {
                const int j = j00 + 16 * g + ri;
                if (j < nout) {
                    const R bj = bias[j];
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int row = Mma<R>::row(kq, r);
                        const R cv = acc[g][r] + bj;
                        c[row * ldc + j] = cv;
                        out[row * ld + j] = act_f(act, cv);
                    }
                }
            }
        }
        }
        __syncthreads();
        wp = bias + nout;
        nin = nout;
        R *tmp = in;
        in = out;
        out = tmp;
    }
    // atomic energy (hdnnps.f90:526); the output layer has one node
    if (lane < tn) Ei_out[atom] = (double)in[lane * ld] + m.einfo[e].atom_energy;

    // backward, layer_backward (:116-158) with gradOut = 1; the gradient with
    // respect to the symmetry functions goes from the accumulators to dEdG
    R *g = out;
    if (lane < T) g[lane * ld] = 1;
    __syncthreads();
#pragma unroll
    for (int l = kMaxLay - 1; l >= 0; --l) {
        if (l >= sh.nlayers) continue;
        const int nout = sh.nodes[l];
        const int ni = l == 0 ? nsf : sh.nodes[l > 0 ? l - 1 : 0];
        const R *c = Cb + sh.coff[l];
        const int ldc = sh.ldc[l];
        const char act = sh.act[l];
        for (int t = lane; t < T * nout; t += 64) {
            const int a = t / nout, o = t - a * nout;
            g[a * ld + o] *= act_d(act, c[a * ldc + o]);
        }
        __syncthreads();
        R *gn = (g == P0) ? P1 : P0;
        const R *wl = Wl[l];
        for (int j00 = 0; j00 < ni; j00 += 16 * kJG<R>) {
            V acc[kJG<R>];
#pragma unroll
            for (int gi = 0; gi < kJG<R>; ++gi) acc[gi] = V{0, 0, 0, 0};
            for (int k0 = 0; k0 < nout; k0 += 4 * kKU) {
                R a[kKU], b[kJG<R>][kKU];
#pragma unroll
                for (int u = 0; u < kKU; ++u) {
                    const int k = k0 + 4 * u + kq;
                    a[u] = k < nout ? g[ri * ld + k] : R(0);
#pragma unroll
                    for (int gi = 0; gi < kJG<R>; ++gi) {
                        const int j = j00 + 16 * gi + ri;
                        b[gi][u] = (k < nout && j < ni) ? wl[(size_t)j * nout + k] : R(0);
                    }
                }
#pragma unroll
                for (int gi = 0; gi < kJG<R>; ++gi)
                    if (j00 + 16 * gi < ni)
#pragma unroll
                        for (int u = 0; u < kKU; ++u) acc[gi] = Mma<R>::op(a[u], b[gi][u], acc[gi]);
            }
#pragma unroll
            for (int gi = 0; gi < kJG<R>; ++gi) {
                const int j = j00 + 16 * gi + ri;
                if (j < ni) {
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int row = Mma<R>::row(kq, r);
                        if (l > 0) {
                            gn[row * ld + j] = acc[gi][r];
                        } else if (row < tn) {
                            R *d = dEdG + (size_t)rows[row] * ldg + j;
                            const R v = scale * acc[gi][r];
                            *d = accumulate ? *d + v : v;
                        }
                    }
                }
            }
        }
        __syncthreads();
        g = gn;
    }
    // the padding columns nsf..nsfmax-1 of the tile's rows
    if (!accumulate) {
        const int pad = ldg - nsf;
        for (int t = lane; t < T * pad; t += 64) {
            const int a = t / pad;
            if (a < tn) dEdG[(size_t)rows[a] * ldg + nsf + (t - a * pad)] = R(0);
        }
    }
}

static int odd(int x) { return x | 1; }

// the two activation buffers hold the hidden layers only (the symmetry
// functions and their gradient stay in global memory)
static int hidden_width(const DevModel &m)
{
    int w = 1;
    for (int l = 0; l < m.nlayers; ++l) w = w > m.nodes[l] ? w : m.nodes[l];
    return w;
}

size_t mlp_lds_bytes(bool fp64, const DevModel &m, int)
{
    size_t tot = 2 * 16 * (size_t)odd(hidden_width(m));
    for (int l = 0; l < m.nlayers; ++l) tot += 16 * (size_t)odd(m.nodes[l]);
    return tot * (fp64 ? 8 : 4) + 16 * sizeof(int);
}

template <typename R>
void launch_mlp_t(const DevModel &m, Work &w, int maxw, int nblocks, const R *G, R *dEdG, double *Ei, R scale,
                  bool accumulate, hipStream_t s, bool bf16)
{
    MlpShape sh{};
    sh.nlayers = m.nlayers;
    sh.ld = odd(hidden_width(m));
    (void)maxw;
    int off = 0;
    for (int l = 0; l < m.nlayers; ++l) {
        sh.nodes[l] = m.nodes[l];
        sh.act[l] = m.act[l];
        sh.ldc[l] = odd(m.nodes[l]);
        sh.coff[l] = off;
        off += 16 * sh.ldc[l];
    }
    sh.total = 2 * 16 * sh.ld + off;
    const size_t lds = sizeof(R) * (size_t)sh.total + 16 * sizeof(int);
    if (bf16 && sizeof(R) == 4)
        hipLaunchKernelGGL((k_mlp<R, true>), dim3(nblocks), dim3(64), lds, s, m, w, sh, G, dEdG, Ei, scale,
                           accumulate ? 1 : 0);
    else
        hipLaunchKernelGGL((k_mlp<R, false>), dim3(nblocks), dim3(64), lds, s, m, w, sh, G, dEdG, Ei, scale,
                           accumulate ? 1 : 0);
}

template void launch_mlp_t<double>(const DevModel &, Work &, int, int, const double *, double *, double *, double,
                                   bool, hipStream_t, bool);
template void launch_mlp_t<float>(const DevModel &, Work &, int, int, const float *, float *, double *, float, bool,
                                  hipStream_t, bool);

}  // namespace hdnnp
