// Symmetry-function kernels: G2/G4 values (pass A) and their gradients
// contracted with dE/dG (pass B). dG/dR is never stored: pass B recomputes the
// pair and triplet terms and folds the SF sum into three scalars per triplet.
//
// Reference: calcAllSymmetryFunctions and cutoff
// (src/potentials/symmetryFunctions.f90:158-327), the scaling of
// hdnnpCalcSymfunctions (src/potentials/hdnnps.f90:471-481) and the force /
// dEdLat scatter of hdnnpCalcNNs (hdnnps.f90:533-553).
//
// One 64-lane wavefront (= one work-group) per atom. The atom's neighbours
// within its cutoff are staged in LDS grouped by element. Triplets (i, j, k)
// whose third side r_jk reaches the cutoff contribute nothing (fc(r_jk) = 0
// and fc'(r_jk) = 0 there), so each pass first compacts the triplets with
// r_jk < rc into an LDS list and evaluates only those, with all 64 lanes busy.
//   values:    per element pair (f, g) of the neighbours, the block's SF
//              parameters are wave-uniform (scalar registers); each lane
//              accumulates its triplets in registers; one DPP reduction per SF.
//              G2: one lane per SF, looping over that SF's neighbour element.
//   gradients: all triplets in circulant order (j, j+d mod n), so the lanes of
//              one round touch distinct j and distinct k; each lane reads its
//              block's SF parameters from an LDS table and adds g_j, g_k to the
//              per-neighbour LDS accumulators without contention.
// The FP32 path uses the hardware exp/rcp/sqrt approximations; the FP64 path
// keeps the libm-accurate functions.
#include <rocprim/intrinsics/warp_shuffle.hpp>

#include "device.hpp"

namespace hdnnp {

constexpr int kPairChunk = 512;    // triplet candidates compacted per LDS chunk

// ---------------------------------------------------------------- math
template <typename R> struct M;
template <> struct M<double> {
    static __device__ double exp(double x) { return ::exp(x); }
    static __device__ double sqrt(double x) { return ::sqrt(x); }
    static __device__ double rcp(double x) { return 1.0 / x; }
    static __device__ double tanh(double x) { return ::tanh(x); }
    static __device__ double pow(double x, double y) { return ::pow(x, y); }
};
template <> struct M<float> {
    static __device__ float exp(float x) { return __expf(x); }
    static __device__ float sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
    static __device__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
    // tanh on [0, 1] through one exponential
    static __device__ float tanh(float x)
    {
        const float e = __expf(-2.0f * x);
        return (1.0f - e) * __builtin_amdgcn_rcpf(1.0f + e);
    }
    static __device__ float pow(float x, float y) { return __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x)); }
};

template <typename R> __device__ __forceinline__ R tab(const double *d, const float *f, int x);
template <> __device__ __forceinline__ double tab<double>(const double *d, const float *, int x) { return d[x]; }
template <> __device__ __forceinline__ float tab<float>(const double *, const float *f, int x) { return f[x]; }

// tanh^3 cutoff, symmetryFunctions.f90:310-327; dfc = fc'(r)/r
template <typename R>
__device__ __forceinline__ void cutoff_fn(R r, R rc, R rrc, R &fc, R &dfc)
{
    if (r < rc) {
        const R t = M<R>::tanh(R(1) - r * rrc);
        const R tt = t * t;
        fc = tt * t;
        dfc = -(R(3) * rrc * M<R>::rcp(r)) * (tt - tt * tt);
    } else {
        fc = 0;
        dfc = 0;
    }
}

// b^(zeta-1) with a wave-uniform exponent: integer powers by repeated
// squaring (the integer branch of symmetryFunctions.f90:265-273)
template <typename R>
__device__ __forceinline__ R zpow_uniform(R b, int ip, R zm1)
{
    if (ip == kNonIntegralZeta) return b > R(0) ? M<R>::pow(b, zm1) : (zm1 > R(0) ? R(0) : R(1));
    int n = ip < 0 ? -ip : ip;
    R r = 1, p = b;
    while (n) {
        if (n & 1) r *= p;
        n >>= 1;
        if (n) p *= p;
    }
    return ip < 0 ? R(1) / r : r;
}

// the same with a per-lane exponent: branch-free over `bits` binary digits
template <typename R>
__device__ __forceinline__ R zpow_lane(R b, int ip, R zm1, int bits)
{
    if (ip == kNonIntegralZeta) return b > R(0) ? M<R>::pow(b, zm1) : (zm1 > R(0) ? R(0) : R(1));
    const int n = ip < 0 ? -ip : ip;
    R r = 1, p = b;
    for (int k = 0; k < bits; ++k) {
        r = ((n >> k) & 1) ? r * p : r;
        p = p * p;
    }
    return ip < 0 ? R(1) / r : r;
}

// full-wave sum with DPP (the result is read from lane 63)
template <typename R>
__device__ __forceinline__ R wave_total(R v)
{
    using rocprim::detail::warp_move_dpp;
    v += warp_move_dpp<R, 0xb1>(v);    // quad_perm [1,0,3,2]
    v += warp_move_dpp<R, 0x4e>(v);    // quad_perm [2,3,0,1]
    v += warp_move_dpp<R, 0x124>(v);   // row_ror:4
    v += warp_move_dpp<R, 0x128>(v);   // row_ror:8
    v += warp_move_dpp<R, 0x142>(v);   // row_bcast:15
    v += warp_move_dpp<R, 0x143>(v);   // row_bcast:31
    return rocprim::warp_shuffle(v, 63, 64);
}

template <typename R>
__device__ __forceinline__ R wave_sum(R v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

__device__ __forceinline__ int mbcnt(unsigned long long m)
{
    return __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

// ---------------------------------------------------------------- LDS layout
// 32-bit pointers into LDS: they stay in scalar registers and keep the LDS
// address space, so accesses compile to ds_* instructions
#define LDS __attribute__((address_space(3)))

template <typename R>
struct Lds {
    LDS R *dx, *dy, *dz, *r, *ir, *fc, *dfc;   // staged neighbours
    LDS R *acc;                                 // values: G sums; gradients: dE/dG of raw G
    LDS double *gx, *gy, *gz;                   // gradients: per-neighbour accumulators (FP64:
                                                // ds_add_f64 is much faster than ds_add_f32 here)
    LDS int *ent, *el;                          // list entry and element rank of a staged neighbour
    LDS int *pairs;                             // compacted triplets
    LDS R *t_eta, *t_lam, *t_zeta, *t_pre;      // gradients: G4 table of the centre element
    LDS int *t_s, *t_ip, *t_new;
    LDS int *btab;                              // gradients: start, count of block f*kMaxEl+g at [2b], [2b+1]
};

template <typename R>
__device__ __forceinline__ void lds_add(LDS R *p, R v)
{
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// LDS layout. Everything sized by the staging capacity nmax comes first, so
// that with nmax a compile-time constant all these arrays sit at constant
// offsets; the per-SF accumulators and the SF table follow.
struct LdsOffsets {
    size_t R, D, I, P, B, A, T, TI, total;
};

template <typename R>
__device__ __host__ __forceinline__ LdsOffsets lds_offsets(int nmax, int nsfmax, int g4max, bool grad)
{
    LdsOffsets o{};
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t at = off;
        off += (bytes + 15) / 16 * 16;
        return at;
    };
    o.R = take(sizeof(R) * (size_t)nmax * 7);
    o.D = grad ? take(sizeof(double) * (size_t)nmax * 3) : 0;
    o.I = take(sizeof(int) * 2 * (size_t)nmax);
    o.P = take(sizeof(int) * kPairChunk);
    o.B = grad ? take(sizeof(int) * 2 * kMaxEl * kMaxEl) : 0;
    o.A = take(sizeof(R) * (size_t)nsfmax);
    o.T = grad ? take(sizeof(R) * 4 * (size_t)g4max) : 0;
    o.TI = grad ? take(sizeof(int) * 3 * (size_t)g4max) : 0;
    o.total = off;
    return o;
}

template <typename R>
__host__ __device__ inline size_t lds_bytes(int nmax, int nsfmax, int g4max, bool grad)
{
    return lds_offsets<R>(nmax, nsfmax, g4max, grad).total;
}

template <typename R>
__device__ __forceinline__ Lds<R> make_lds(int nmax, int nsfmax, int g4max, bool grad, LDS char *base)
{
    const LdsOffsets o = lds_offsets<R>(nmax, nsfmax, g4max, grad);
    Lds<R> L{};
    LDS R *p = (LDS R *)(base + o.R);
    L.dx = p;
    L.dy = p + nmax;
    L.dz = p + 2 * nmax;
    L.r = p + 3 * nmax;
    L.ir = p + 4 * nmax;
    L.fc = p + 5 * nmax;
    L.dfc = p + 6 * nmax;
    if (grad) {
        LDS double *pd = (LDS double *)(base + o.D);
        L.gx = pd;
        L.gy = pd + nmax;
        L.gz = pd + 2 * nmax;
    }
    L.ent = (LDS int *)(base + o.I);
    L.el = L.ent + nmax;
    L.pairs = (LDS int *)(base + o.P);
    L.acc = (LDS R *)(base + o.A);
    if (grad) {
        L.btab = (LDS int *)(base + o.B);
        LDS R *t = (LDS R *)(base + o.T);
        L.t_eta = t;
        L.t_lam = t + g4max;
        L.t_zeta = t + 2 * g4max;
        L.t_pre = t + 3 * g4max;
        LDS int *ti = (LDS int *)(base + o.TI);
        L.t_s = ti;
        L.t_ip = ti + g4max;
        L.t_new = ti + 2 * g4max;
    }
    return L;
}

// Geometry of one list entry: d = R_i - (R_j + L.h) (FP64), r = |d|.
__device__ __forceinline__ void entry_geometry(const Geo &g, const Work &w, const double4 &pi, int e,
                                      double &d0, double &d1, double &d2, double &r, int &j)
{
    j = w.ent_j[e];
    int sx, sy, sz;
    unpack_shift(w.ent_shift[e], sx, sy, sz);
    const double l0 = ((g.lat[6] * sz) + g.lat[3] * sy) + g.lat[0] * sx;
    const double l1 = ((g.lat[7] * sz) + g.lat[4] * sy) + g.lat[1] * sx;
    const double l2 = ((g.lat[8] * sz) + g.lat[5] * sy) + g.lat[2] * sx;
    const double4 pj = w.wpos[j];
    d0 = pi.x - (pj.x + l0);
    d1 = pi.y - (pj.y + l1);
    d2 = pi.z - (pj.z + l2);
    r = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
}

// Stage the neighbours of atom i with r < rc(e) into LDS, grouped by element
// (seg[f]..seg[f+1]), list order within an element. Entries beyond the cutoff
// contribute nothing (fc = dfc = 0 there); their gradient rows are zeroed.
template <typename R>
__device__ __forceinline__ void stage(const DevModel &m, const Geo &g, const Work &w, int i, int e, const Lds<R> L,
                      int *seg, R *gent)
{
    const int lane = lane_id();
    const int off = i * w.cap, n = w.count[i];
    const double4 pi = w.wpos[i];
    const double rce = m.einfo[e].rc;
    const R rrc = R(1) / (R)rce;
    int cnt[kMaxEl];
#pragma unroll
    for (int f = 0; f < kMaxEl; ++f) cnt[f] = 0;
    for (int b = 0; b < n; b += 64) {
        const int q = b + lane;
        bool keep = false;
        int f = -1;
        if (q < n) {
            double d0, d1, d2, r;
            int j;
            entry_geometry(g, w, pi, off + q, d0, d1, d2, r, j);
            keep = r < rce;
            f = w.eidx[j];
        }
#pragma unroll
        for (int ff = 0; ff < kMaxEl; ++ff)
            if (ff < m.nel) cnt[ff] += __popcll(__ballot(keep && f == ff));
    }
    int start[kMaxEl], run[kMaxEl];
    int acc = 0;
#pragma unroll
    for (int f = 0; f < kMaxEl; ++f) {
        start[f] = acc;
        run[f] = 0;
        if (lane == 0) seg[f] = acc;
        acc += (f < m.nel) ? cnt[f] : 0;
    }
    if (lane == 0) seg[kMaxEl] = acc;
    for (int b = 0; b < n; b += 64) {
        const int q = b + lane;
        bool keep = false;
        int f = -1;
        double d0 = 0, d1 = 0, d2 = 0, r = 1;
        if (q < n) {
            int j;
            entry_geometry(g, w, pi, off + q, d0, d1, d2, r, j);
            keep = r < rce;
            f = w.eidx[j];
        }
        int pos = -1;
#pragma unroll
        for (int ff = 0; ff < kMaxEl; ++ff) {
            if (ff < m.nel) {
                const unsigned long long mk = __ballot(keep && f == ff);
                if (keep && f == ff) pos = start[ff] + run[ff] + mbcnt(mk);
                run[ff] += __popcll(mk);
            }
        }
        if (keep) {
            const R rr = (R)r;
            R fc, dfc;
            cutoff_fn<R>(rr, (R)rce, rrc, fc, dfc);
            L.dx[pos] = (R)d0;
            L.dy[pos] = (R)d1;
            L.dz[pos] = (R)d2;
            L.r[pos] = rr;
            L.ir[pos] = (R)(1.0 / r);
            L.fc[pos] = fc;
            L.dfc[pos] = dfc;
            L.ent[pos] = off + q;
            L.el[pos] = f;
        } else if (q < n && gent) {
            gent[3 * (size_t)(off + q) + 0] = 0;
            gent[3 * (size_t)(off + q) + 1] = 0;
            gent[3 * (size_t)(off + q) + 2] = 0;
        }
    }
}

// candidate p of an element-pair block -> staged indices (j, k), no division
__device__ __forceinline__ void block_pair(int p, bool same, int nf, int ng, float inv_ng, int sf, int sg, int &j, int &k)
{
    if (same) {
        // pairs (a, b), 0 <= a < b < nf, row-major; T(a) = a*(2nf-a-1)/2
        const float c = (float)(2 * nf - 1);
        int a = (int)((c - sqrtf(c * c - 8.0f * (float)p)) * 0.5f);
        a = max(0, min(a, nf - 2));
        if (a > 0 && a * (2 * nf - a - 1) / 2 > p) --a;
        if (a + 1 <= nf - 2 && (a + 1) * (2 * nf - a - 2) / 2 <= p) ++a;
        j = sf + a;
        k = sf + p - a * (2 * nf - a - 1) / 2 + a + 1;
    } else {
        int a = (int)(((float)p + 0.5f) * inv_ng);
        if (a * ng > p) --a;
        if ((a + 1) * ng <= p) ++a;
        j = sf + a;
        k = sg + (p - a * ng);
    }
}

// candidate p of the circulant enumeration of the unordered pairs of n
// neighbours: rows d = 1..(n-1)/2 hold (j, j+d mod n) for every j; for even
// n the row d = n/2 holds (j, j+n/2) for j < n/2.
__device__ __forceinline__ void circulant_pair(int p, int n, float inv_n, int &j, int &k)
{
    const int rows = (n - 1) / 2;
    int d, jj;
    if (p < rows * n) {
        int q = (int)(((float)p + 0.5f) * inv_n);
        if (q * n > p) --q;
        if ((q + 1) * n <= p) ++q;
        d = q + 1;
        jj = p - q * n;
    } else {
        d = n / 2;
        jj = p - rows * n;
    }
    j = jj;
    k = jj + d;
    if (k >= n) k -= n;
}

// ---------------------------------------------------------------- pass A
template <typename R, int NMAX>
__global__ __launch_bounds__(64) void k_sf_values(DevModel m, Geo g, Work w, R *G)
{
    extern __shared__ __align__(16) char smem[];
    __shared__ int seg[kMaxEl + 1];
    const int i = blockIdx.x;
    const int lane = lane_id();
    const int e = w.eidx[i];
    const int nsf = m.einfo[e].nsf;
    const Lds<R> L = make_lds<R>(NMAX, m.nsfmax, 0, false, (LDS char *)smem);
    for (int s = lane; s < nsf; s += 64) L.acc[s] = 0;
    stage<R>(m, g, w, i, e, L, seg, (R *)nullptr);
    __syncthreads();
    const R rce = (R)m.einfo[e].rc;
    const R rce2 = rce * rce, rrc = R(1) / rce;

    // G2, symmetryFunctions.f90:220-230: one lane per SF of the centre element
    for (int t = lane; t < m.einfo[e].g2_count; t += 64) {
        const int x = m.einfo[e].g2_start + t;
        const int f = m.g2_f[x];
        const R eta = tab<R>(m.g2_eta, m.g2f_eta, x), rs = tab<R>(m.g2_rs, m.g2f_rs, x);
        R sum = 0;
        for (int q = seg[f]; q < seg[f + 1]; ++q) {
            const R dr = L.r[q] - rs;
            sum += M<R>::exp(-eta * dr * dr) * L.fc[q];
        }
        L.acc[m.g2_s[x]] += sum;
    }
    // G4, symmetryFunctions.f90:231-305, one element-pair block at a time
    for (int f = 0; f < m.nel; ++f) {
        for (int h = f; h < m.nel; ++h) {
            const int2 tb = m.g4_tab[(e * kMaxEl + f) * kMaxEl + h];
            if (tb.y == 0) continue;
            const int nf = seg[f + 1] - seg[f], nh = seg[h + 1] - seg[h];
            const bool same = f == h;
            const int np = same ? nf * (nf - 1) / 2 : nf * nh;
            if (np == 0) continue;
            const float inv_nh = 1.0f / (float)max(nh, 1);
            for (int c0 = 0; c0 < tb.y; c0 += 8) {
                const int nc = min(8, tb.y - c0);
                R acc[8];
#pragma unroll
                for (int t = 0; t < 8; ++t) acc[t] = 0;
                for (int cb = 0; cb < np; cb += kPairChunk) {
                    const int ce = min(np, cb + kPairChunk);
                    int nv = 0;
                    for (int p0 = cb; p0 < ce; p0 += 64) {      // compact r_jk < rc
                        const int p = p0 + lane;
                        bool ok = false;
                        int j = 0, k = 0;
                        if (p < ce) {
                            block_pair(p, same, nf, nh, inv_nh, seg[f], seg[h], j, k);
                            const R ex = L.dx[k] - L.dx[j], ey = L.dy[k] - L.dy[j], ez = L.dz[k] - L.dz[j];
                            ok = ex * ex + ey * ey + ez * ez < rce2;
                        }
                        const unsigned long long mk = __ballot(ok);
                        if (ok) L.pairs[nv + mbcnt(mk)] = j | (k << 16);
                        nv += __popcll(mk);
                    }
                    __syncthreads();
                    for (int q = lane; q < nv; q += 64) {
                        const int pk = L.pairs[q];
                        const int j = pk & 0xFFFF, k = pk >> 16;
                        const R xj = L.dx[j], yj = L.dy[j], zj = L.dz[j];
                        const R xk = L.dx[k], yk = L.dy[k], zk = L.dz[k];
                        const R ex = xk - xj, ey = yk - yj, ez = zk - zj;   // d_jk = d_ik - d_ij
                        const R rjk2 = ex * ex + ey * ey + ez * ez;
                        const R rjk = M<R>::sqrt(rjk2);
                        R fcjk, dfcjk;
                        cutoff_fn<R>(rjk, rce, rrc, fcjk, dfcjk);
                        const R rij = L.r[j], rik = L.r[k];
                        const R cs = (xj * xk + yj * yk + zj * zk) * L.ir[j] * L.ir[k];
                        const R F = L.fc[j] * L.fc[k] * fcjk;
                        const R r2 = rij * rij + rik * rik + rjk2;
                        R E = 0;
#pragma unroll
                        for (int t = 0; t < 8; ++t)
                            if (t < nc) {
                                const int x = tb.x + c0 + t;
                                if (t == 0 || m.g4_neweta[x]) E = M<R>::exp(-tab<R>(m.g4_eta, m.g4f_eta, x) * r2);
                                const R b = R(1) + tab<R>(m.g4_lam, m.g4f_lam, x) * cs;
                                const R C = b * zpow_uniform<R>(b, m.g4_ipow[x], tab<R>(m.g4_zeta, m.g4f_zeta, x) - R(1));
                                acc[t] += tab<R>(m.g4_pre, m.g4f_pre, x) * C * E * F;
                            }
                    }
                    __syncthreads();
                }
#pragma unroll
                for (int t = 0; t < 8; ++t)
                    if (t < nc) {
                        const R v = wave_total(acc[t]);
                        if (lane == 0) L.acc[m.g4_s[tb.x + c0 + t]] += v;
                    }
            }
        }
    }
    __syncthreads();
    // centre and scale, hdnnps.f90:471-481
    const int so = m.einfo[e].sf_off;
    for (int s = lane; s < m.nsfmax; s += 64)
        G[(size_t)i * m.nsfmax + s] =
            s < nsf ? (R)(((double)L.acc[s] - m.sf_mean[so + s]) / m.sf_den[so + s]) : R(0);
}

// ---------------------------------------------------------------- pass B
template <typename R, int NMAX>
__global__ __launch_bounds__(64) void k_sf_grad(DevModel m, Geo g, Work w, const R *dEdG,
                                                R *gent)
{
    extern __shared__ __align__(16) char smem[];
    __shared__ int seg[kMaxEl + 1];
    const int i = blockIdx.x;
    const int lane = lane_id();
    const int e = w.eidx[i];
    const int nsf = m.einfo[e].nsf;
    const int so = m.einfo[e].sf_off;
    const Lds<R> L = make_lds<R>(NMAX, m.nsfmax, m.g4_max, true, (LDS char *)smem);
    // dE/dG of the unscaled G: the scaled G is (G - mean)/den
    for (int s = lane; s < nsf; s += 64)
        L.acc[s] = (R)((double)dEdG[(size_t)i * m.nsfmax + s] / m.sf_den[so + s]);
    // G4 table of the centre element; block (f,g) -> (start, count) relative
    const int g4s = m.einfo[e].g4_start, g4n = m.einfo[e].g4_count;
    for (int t = lane; t < g4n; t += 64) {
        const int x = g4s + t;
        L.t_eta[t] = tab<R>(m.g4_eta, m.g4f_eta, x);
        L.t_lam[t] = tab<R>(m.g4_lam, m.g4f_lam, x);
        L.t_zeta[t] = tab<R>(m.g4_zeta, m.g4f_zeta, x);
        L.t_pre[t] = tab<R>(m.g4_pre, m.g4f_pre, x);
        L.t_s[t] = m.g4_s[x];
        L.t_ip[t] = m.g4_ipow[x];
        L.t_new[t] = m.g4_neweta[x];
    }
    for (int t = lane; t < m.nel * kMaxEl; t += 64) {
        const int f = t / kMaxEl, h = t - f * kMaxEl;
        int2 tb = make_int2(0, 0);
        if (h < m.nel) {
            tb = m.g4_tab[(e * kMaxEl + min(f, h)) * kMaxEl + max(f, h)];
            tb.x -= g4s;
        }
        L.btab[2 * t] = tb.x;
        L.btab[2 * t + 1] = tb.y;
    }
    stage<R>(m, g, w, i, e, L, seg, gent);
    __syncthreads();
    const int n = seg[kMaxEl];
    for (int q = lane; q < n; q += 64) L.gx[q] = L.gy[q] = L.gz[q] = 0;
    __syncthreads();
    const R rce = (R)m.einfo[e].rc;
    const R rce2 = rce * rce, rrc = R(1) / rce;

    // G2: g_j = d_ij * (-dfc * sum w E + fc/r * sum w 2 eta (r - Rs) E), one lane per neighbour
    for (int q = lane; q < n; q += 64) {
        const int2 tb = m.g2_tab[e * kMaxEl + L.el[q]];
        const R rr = L.r[q];
        R A = 0, B = 0;
        for (int t = 0; t < tb.y; ++t) {
            const int x = tb.x + t;
            const R eta = tab<R>(m.g2_eta, m.g2f_eta, x);
            const R dr = rr - tab<R>(m.g2_rs, m.g2f_rs, x);
            const R ws = L.acc[m.g2_s[x]] * M<R>::exp(-eta * dr * dr);
            A += ws;
            B += ws * R(2) * eta * dr;
        }
        const R c = -L.dfc[q] * A + L.fc[q] * B * L.ir[q];
        L.gx[q] += (double)(c * L.dx[q]);
        L.gy[q] += (double)(c * L.dy[q]);
        L.gz[q] += (double)(c * L.dz[q]);
    }
    __syncthreads();
    // G4: per triplet, S1 = sum w P C E, S2 = sum w P C E 2 eta,
    // S3 = sum w P b^(zeta-1) lambda zeta E, then
    // g_j = S1 dF/dR_j + S2 F (d_ij - d_jk) + S3 F dcos/dR_j, same for k.
    const int np = n * (n - 1) / 2;
    const float inv_n = 1.0f / (float)max(n, 1);
    for (int cb = 0; cb < np; cb += kPairChunk) {
        const int ce = min(np, cb + kPairChunk);
        int nv = 0;
        for (int p0 = cb; p0 < ce; p0 += 64) {                  // compact r_jk < rc
            const int p = p0 + lane;
            bool ok = false;
            int pk = 0;
            if (p < ce) {
                int j, k;
                circulant_pair(p, n, inv_n, j, k);
                const int ej = L.el[j], ek = L.el[k];
                const int blk = min(ej, ek) * kMaxEl + max(ej, ek);
                const R ex = L.dx[k] - L.dx[j], ey = L.dy[k] - L.dy[j], ez = L.dz[k] - L.dz[j];
                ok = L.btab[2 * blk + 1] > 0 && ex * ex + ey * ey + ez * ez < rce2;
                pk = j | (k << 10) | (blk << 20);
            }
            const unsigned long long mk = __ballot(ok);
            if (ok) L.pairs[nv + mbcnt(mk)] = pk;
            nv += __popcll(mk);
        }
        __syncthreads();
        for (int q = lane; q < nv; q += 64) {
            const int pk = L.pairs[q];
            const int j = pk & 0x3FF, k = (pk >> 10) & 0x3FF;
            const int tbx = L.btab[2 * (pk >> 20)], tby = L.btab[2 * (pk >> 20) + 1];
            const R xj = L.dx[j], yj = L.dy[j], zj = L.dz[j];
            const R xk = L.dx[k], yk = L.dy[k], zk = L.dz[k];
            const R ex = xk - xj, ey = yk - yj, ez = zk - zj;
            const R rjk2 = ex * ex + ey * ey + ez * ez;
            const R rjk = M<R>::sqrt(rjk2);
            R fcjk, dfcjk;
            cutoff_fn<R>(rjk, rce, rrc, fcjk, dfcjk);
            const R rij = L.r[j], rik = L.r[k], irj = L.ir[j], irk = L.ir[k];
            const R fcij = L.fc[j], fcik = L.fc[k], dfcij = L.dfc[j], dfcik = L.dfc[k];
            const R inv = irj * irk;
            const R cs = (xj * xk + yj * yk + zj * zk) * inv;
            const R F = fcij * fcik * fcjk;
            const R r2 = rij * rij + rik * rik + rjk2;
            R S1 = 0, S2 = 0, S3 = 0, E = 0;
            for (int t = tbx; t < tbx + tby; ++t) {
                const R eta = L.t_eta[t];
                if (t == tbx || L.t_new[t]) E = M<R>::exp(-eta * r2);
                const R lam = L.t_lam[t];
                const R zeta = L.t_zeta[t];
                const R b = R(1) + lam * cs;
                const R Bp = zpow_lane<R>(b, L.t_ip[t], zeta - R(1), m.powbits);
                const R wpe = L.acc[L.t_s[t]] * L.t_pre[t] * E;
                const R wpc = wpe * Bp * b;
                S1 += wpc;
                S2 += wpc * R(2) * eta;
                S3 += wpe * Bp * lam * zeta;
            }
            const R aj = fcik * dfcjk * fcij, bj = fcik * dfcij * fcjk;   // dF/dR_j = aj d_jk - bj d_ij
            const R ak = fcij * dfcjk * fcik, bk = fcij * dfcik * fcjk;   // dF/dR_k = -(ak d_jk + bk d_ik)
            const R cj = cs * irj * irj, ck = cs * irk * irk;
            const R S2F = S2 * F, S3F = S3 * F;
            const R gjx = S1 * (aj * ex - bj * xj) + S2F * (xj - ex) + S3F * (-xk * inv + xj * cj);
            const R gjy = S1 * (aj * ey - bj * yj) + S2F * (yj - ey) + S3F * (-yk * inv + yj * cj);
            const R gjz = S1 * (aj * ez - bj * zj) + S2F * (zj - ez) + S3F * (-zk * inv + zj * cj);
            const R gkx = -S1 * (ak * ex + bk * xk) + S2F * (xk + ex) + S3F * (-xj * inv + xk * ck);
            const R gky = -S1 * (ak * ey + bk * yk) + S2F * (yk + ey) + S3F * (-yj * inv + yk * ck);
            const R gkz = -S1 * (ak * ez + bk * zk) + S2F * (zk + ez) + S3F * (-zj * inv + zk * ck);
            lds_add<double>(&L.gx[j], (double)gjx);
            lds_add<double>(&L.gy[j], (double)gjy);
            lds_add<double>(&L.gz[j], (double)gjz);
            lds_add<double>(&L.gx[k], (double)gkx);
            lds_add<double>(&L.gy[k], (double)gky);
            lds_add<double>(&L.gz[k], (double)gkz);
        }
        __syncthreads();
    }
    // per-entry gradient rows and the atom's virial -sum_n g_n (x) d_in
    double v[9];
#pragma unroll
    for (int t = 0; t < 9; ++t) v[t] = 0;
    for (int q = lane; q < n; q += 64) {
        const double gx = L.gx[q], gy = L.gy[q], gz = L.gz[q];
        const size_t en = (size_t)L.ent[q];
        gent[3 * en + 0] = (R)gx;
        gent[3 * en + 1] = (R)gy;
        gent[3 * en + 2] = (R)gz;
        const double dxq = L.dx[q], dyq = L.dy[q], dzq = L.dz[q];
        v[0] -= gx * dxq; v[1] -= gx * dyq; v[2] -= gx * dzq;
        v[3] -= gy * dxq; v[4] -= gy * dyq; v[5] -= gy * dzq;
        v[6] -= gz * dxq; v[7] -= gz * dyq; v[8] -= gz * dzq;
    }
#pragma unroll
    for (int t = 0; t < 9; ++t) {
        const double s = wave_total(v[t]);
        if (lane == 0) w.vir[9 * (size_t)i + t] = s;
    }
}

template <typename R>
void launch_sf_values(const DevModel &m, const Geo &g, Work &w, int nmax, R *G, hipStream_t s)
{
#define HDNNP_SFV(N)                                                                          \
    hipLaunchKernelGGL((k_sf_values<R, N>), dim3(w.nat), dim3(64),                         \
                       lds_bytes<R>(N, m.nsfmax, 0, false), s, m, g, w, G)
    if (nmax <= 64) HDNNP_SFV(64);
    else if (nmax <= 128) HDNNP_SFV(128);
    else if (nmax <= 256) HDNNP_SFV(256);
    else if (nmax <= 512) HDNNP_SFV(512);
    else HDNNP_SFV(1024);
#undef HDNNP_SFV
}

template <typename R>
void launch_sf_grad(const DevModel &m, const Geo &g, Work &w, int nmax, const R *dEdG, R *gent,
                    hipStream_t s)
{
#define HDNNP_SFG(N)                                                                          \
    hipLaunchKernelGGL((k_sf_grad<R, N>), dim3(w.nat), dim3(64),                           \
                       lds_bytes<R>(N, m.nsfmax, m.g4_max, true), s, m, g, w, dEdG, gent)
    if (nmax <= 64) HDNNP_SFG(64);
    else if (nmax <= 128) HDNNP_SFG(128);
    else if (nmax <= 256) HDNNP_SFG(256);
    else if (nmax <= 512) HDNNP_SFG(512);
    else HDNNP_SFG(1024);
#undef HDNNP_SFG
}

template void launch_sf_values<double>(const DevModel &, const Geo &, Work &, int, double *, hipStream_t);
template void launch_sf_values<float>(const DevModel &, const Geo &, Work &, int, float *, hipStream_t);
template void launch_sf_grad<double>(const DevModel &, const Geo &, Work &, int, const double *, double *,
                                     hipStream_t);
template void launch_sf_grad<float>(const DevModel &, const Geo &, Work &, int, const float *, float *,
                                    hipStream_t);

// staging capacity of the kernel instance that serves nmax neighbours
int sf_bucket(int nmax)
{
    int b = 64;
    while (b < nmax) b *= 2;
    return b;
}

size_t sf_lds_bytes(bool fp64, int nmax, int nsfmax, int g4max, bool grad)
{
    nmax = sf_bucket(nmax);
    return fp64 ? lds_bytes<double>(nmax, nsfmax, g4max, grad) : lds_bytes<float>(nmax, nsfmax, g4max, grad);
}

}  // namespace hdnnp
