// Per-element MLP: forward pass and backpropagation of dE/dG.
//
// Reference: nn_forward / layer_forward and nn_backward / layer_backward
// (src/neuralnetwork/neuralNetworks.f90:94-158, 220-313) with activations
// activate / deactivate (:33-84), called per atom by hdnnpCalcNNs
// (src/potentials/hdnnps.f90:523-531) with gradOut = 1.
//
// One work-group runs a tile of T atoms of one element (atoms grouped by
// element through `perm`). Layer activations of the tile live in LDS: the
// pre-activations c_l of every layer (needed by the backward pass) and two
// ping-pong buffers for the running activation / gradient.
#include "device.hpp"

namespace hdnnp {

__device__ inline double act_f(char a, double x)
{
    switch (a) {
    case 't': return tanh(x);
    case 'p': return log(1.0 + exp(x));
    case 'r': return x > 0.0 ? x : 0.0;
    default: return x;
    }
}
__device__ inline double act_d(char a, double x)
{
    switch (a) {
    case 't': { const double t = tanh(x); return 1.0 - t * t; }
    case 'p': return 1.0 / (1.0 + exp(-x));
    case 'r': return x > 0.0 ? 1.0 : 0.0;
    default: return 1.0;
    }
}
__device__ inline float act_f(char a, float x)
{
    switch (a) {
    case 't': return tanhf(x);
    case 'p': return logf(1.0f + __expf(x));
    case 'r': return x > 0.0f ? x : 0.0f;
    default: return x;
    }
}
__device__ inline float act_d(char a, float x)
{
    switch (a) {
    case 't': { const float t = tanhf(x); return 1.0f - t * t; }
    case 'p': return 1.0f / (1.0f + __expf(-x));
    case 'r': return x > 0.0f ? 1.0f : 0.0f;
    default: return 1.0f;
    }
}

template <typename R> __device__ inline const R *weights(const DevModel &m, int e);
template <> __device__ inline const double *weights<double>(const DevModel &m, int e) { return m.W64[e]; }
template <> __device__ inline const float *weights<float>(const DevModel &m, int e) { return m.W32[e]; }

template <typename R>
__global__ __launch_bounds__(256) void k_mlp(DevModel m, Work w, int T, int maxw, const R *G, R *dEdG)
{
    extern __shared__ __align__(16) char smem[];
    // which (element, tile) this work-group runs
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
    const int nsf = m.nsf[e];
    const int ld = m.nsfmax;
    int sumc = 0;
    for (int l = 0; l < m.nlayers; ++l) sumc += m.nodes[l];
    R *cbuf = (R *)smem;                 // [sum nodes][T]: c_l stored at layer offset
    R *P0 = cbuf + (size_t)sumc * T;     // [T][maxw]
    R *P1 = P0 + (size_t)T * maxw;
    const R *W = weights<R>(m, e);

    for (int t = threadIdx.x; t < T * nsf; t += blockDim.x) {
        const int a = t / nsf, s = t - a * nsf;
        P0[a * maxw + s] = a < tn ? G[(size_t)w.perm[t0 + a] * ld + s] : R(0);
    }
    __syncthreads();

    // forward, layer_forward (neuralNetworks.f90:94-114): c = W x + b, y = act(c)
    const R *Wl[kMaxLay];
    int nin = nsf, coff = 0;
    R *in = P0, *out = P1;
    const R *wp = W;
    for (int l = 0; l < m.nlayers; ++l) {
        const int nout = m.nodes[l];
        Wl[l] = wp;
        const R *bias = wp + (size_t)nin * nout;
        R *c = cbuf + (size_t)coff * T;
        const char act = m.act[l];
        for (int t = threadIdx.x; t < T * nout; t += blockDim.x) {
            const int a = t / nout, o = t - a * nout;
            R acc = 0;
            for (int j = 0; j < nin; ++j) acc += wp[j * nout + o] * in[a * maxw + j];
            const R cv = acc + bias[o];
            c[a * nout + o] = cv;
            out[a * maxw + o] = act_f(act, cv);
        }
        __syncthreads();
        wp = bias + nout;
        coff += nout;
        nin = nout;
        R *tmp = in;
        in = out;
        out = tmp;
    }
    // atomic energy (hdnnps.f90:526); the output layer has one node
    for (int a = threadIdx.x; a < tn; a += blockDim.x)
        w.Ei[w.perm[t0 + a]] = (double)in[a * maxw] + m.atom_energy[e];
    __syncthreads();

    // backward, layer_backward (:116-158) with gradOut = 1
    R *g = out;
    for (int a = threadIdx.x; a < T; a += blockDim.x) g[a * maxw] = 1;
    __syncthreads();
    for (int l = m.nlayers - 1; l >= 0; --l) {
        const int nout = m.nodes[l];
        const int ni = l == 0 ? nsf : m.nodes[l - 1];
        coff -= nout;
        const R *c = cbuf + (size_t)coff * T;
        const char act = m.act[l];
        for (int t = threadIdx.x; t < T * nout; t += blockDim.x) {
            const int a = t / nout, o = t - a * nout;
            g[a * maxw + o] *= act_d(act, c[a * nout + o]);
        }
        __syncthreads();
        R *gn = (g == P0) ? P1 : P0;
        const R *wl = Wl[l];
        for (int t = threadIdx.x; t < T * ni; t += blockDim.x) {
            const int a = t / ni, j = t - a * ni;
            R acc = 0;
            for (int o = 0; o < nout; ++o) acc += g[a * maxw + o] * wl[j * nout + o];
            gn[a * maxw + j] = acc;
        }
        __syncthreads();
        g = gn;
    }
    for (int t = threadIdx.x; t < T * ld; t += blockDim.x) {
        const int a = t / ld, s = t - a * ld;
        if (a < tn) dEdG[(size_t)w.perm[t0 + a] * ld + s] = s < nsf ? g[a * maxw + s] : R(0);
    }
}

size_t mlp_lds_bytes(bool fp64, int T, int sumc, int maxw)
{
    return (fp64 ? 8 : 4) * ((size_t)sumc * T + 2 * (size_t)T * maxw);
}

template <typename R>
void launch_mlp_t(const DevModel &m, Work &w, int T, int maxw, int nblocks, const R *G, R *dEdG,
                  hipStream_t s)
{
    int sumc = 0;
    for (int l = 0; l < m.nlayers; ++l) sumc += m.nodes[l];
    const size_t sh = mlp_lds_bytes(sizeof(R) == 8, T, sumc, maxw);
    hipLaunchKernelGGL(k_mlp<R>, dim3(nblocks), dim3(256), sh, s, m, w, T, maxw, G, dEdG);
}

template void launch_mlp_t<double>(const DevModel &, Work &, int, int, int, const double *, double *,
                                   hipStream_t);
template void launch_mlp_t<float>(const DevModel &, Work &, int, int, int, const float *, float *,
                                  hipStream_t);

}  // namespace hdnnp
