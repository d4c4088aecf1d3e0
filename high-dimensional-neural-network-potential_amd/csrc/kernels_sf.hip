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
// within its cutoff are staged in LDS grouped by element; G2 terms run with
// lanes over neighbours, G4 terms with lanes over the neighbour pairs of one
// element pair, so every lane of a pass evaluates the same symmetry functions.
#include "device.hpp"

namespace hdnnp {

__device__ inline double f_exp(double x) { return exp(x); }
__device__ inline float f_exp(float x) { return __expf(x); }
__device__ inline double f_tanh(double x) { return tanh(x); }
__device__ inline float f_tanh(float x) { return tanhf(x); }
__device__ inline double f_sqrt(double x) { return sqrt(x); }
__device__ inline float f_sqrt(float x) { return sqrtf(x); }
__device__ inline double f_pow(double x, double y) { return pow(x, y); }
__device__ inline float f_pow(float x, float y) { return powf(x, y); }

template <typename R>
__device__ inline R wave_sum(R v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

template <typename R>
__device__ inline R ipow(R b, int n)
{
    R r = 1;
    while (n) {
        if (n & 1) r *= b;
        n >>= 1;
        if (n) b *= b;
    }
    return r;
}

// b^(zeta-1): integer powers by repeated squaring (symmetryFunctions.f90:
// 265-273 take the integer branch when |zeta - int(zeta)| < 1e-10)
template <typename R>
__device__ inline R zpow(R b, int ip, R zeta)
{
    if (ip != kNonIntegralZeta) return ip >= 0 ? ipow<R>(b, ip) : R(1) / ipow<R>(b, -ip);
    return f_pow(fmax(b, R(0)), zeta - R(1));
}

// tanh^3 cutoff, symmetryFunctions.f90:310-327; dfc = fc'(r)/r
template <typename R>
__device__ inline void cutoff_fn(R r, R rc, R &fc, R &dfc)
{
    if (r < rc) {
        const R t = f_tanh(R(1) - r / rc);
        const R tt = t * t;
        fc = tt * t;
        dfc = -(R(3) / (rc * r)) * (tt - tt * tt);
    } else {
        fc = 0;
        dfc = 0;
    }
}

__device__ inline int lane_id() { return threadIdx.x & 63; }

__device__ inline int mbcnt(unsigned long long m)
{
    return __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

template <typename R>
struct Lds {
    R *dx, *dy, *dz, *r, *fc, *dfc;
    R *acc;             // per-SF accumulator (values) / dE/dG of raw G (gradients)
    R *gx, *gy, *gz;    // per-neighbour gradient accumulators
    int *ent;           // entry index of each staged neighbour
};

template <typename R>
__device__ inline Lds<R> carve(char *smem, int nmax, int nsfmax, bool grad)
{
    Lds<R> L;
    R *p = (R *)smem;
    L.dx = p; p += nmax;
    L.dy = p; p += nmax;
    L.dz = p; p += nmax;
    L.r = p; p += nmax;
    L.fc = p; p += nmax;
    L.dfc = p; p += nmax;
    L.acc = p; p += nsfmax;
    if (grad) {
        L.gx = p; p += nmax;
        L.gy = p; p += nmax;
        L.gz = p; p += nmax;
    } else {
        L.gx = L.gy = L.gz = nullptr;
    }
    L.ent = (int *)p;
    return L;
}

template <typename R>
size_t lds_bytes(int nmax, int nsfmax, bool grad)
{
    return sizeof(R) * ((size_t)nmax * (grad ? 9 : 6) + nsfmax) + sizeof(int) * nmax;
}

// Geometry of one list entry: d = R_i - (R_j + L.h) (FP64), r = |d|.
__device__ inline void entry_geometry(const Geo &g, const Work &w, const double4 &pi, int e,
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
__device__ void stage(const DevModel &m, const Geo &g, const Work &w, int i, int e, Lds<R> &L,
                      int *seg, R *gent)
{
    const int lane = lane_id();
    const int off = w.offset[i], n = w.count[i];
    const double4 pi = w.wpos[i];
    const double rce = m.rc[e];
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
            cutoff_fn<R>(rr, (R)rce, fc, dfc);
            L.dx[pos] = (R)d0;
            L.dy[pos] = (R)d1;
            L.dz[pos] = (R)d2;
            L.r[pos] = rr;
            L.fc[pos] = fc;
            L.dfc[pos] = dfc;
            L.ent[pos] = off + q;
        } else if (q < n && gent) {
            gent[3 * (size_t)(off + q) + 0] = 0;
            gent[3 * (size_t)(off + q) + 1] = 0;
            gent[3 * (size_t)(off + q) + 2] = 0;
        }
    }
}

// pair p of an element-pair block -> staged indices (j, k)
__device__ inline void decode_pair(int p, bool same, int nf, int ng, int sf, int sg, int &j, int &k)
{
    if (same) {
        // pairs (a, b), 0 <= a < b < nf, row-major; T(a) = a*(2nf-a-1)/2
        const float c = (float)(2 * nf - 1);
        int a = (int)((c - sqrtf(c * c - 8.0f * (float)p)) * 0.5f);
        a = max(0, min(a, nf - 2));
        while (a > 0 && a * (2 * nf - a - 1) / 2 > p) --a;
        while (a + 1 <= nf - 2 && (a + 1) * (2 * nf - a - 2) / 2 <= p) ++a;
        const int b = p - a * (2 * nf - a - 1) / 2 + a + 1;
        j = sf + a;
        k = sf + b;
    } else {
        const int a = p / ng;
        j = sf + a;
        k = sg + (p - a * ng);
    }
}

template <typename R>
__global__ __launch_bounds__(64) void k_sf_values(DevModel m, Geo g, Work w, int nmax, R *G)
{
    extern __shared__ __align__(16) char smem[];
    __shared__ int seg[kMaxEl + 1];
    const int i = blockIdx.x;
    const int lane = lane_id();
    const int e = w.eidx[i];
    const int nsf = m.nsf[e];
    Lds<R> L = carve<R>(smem, nmax, m.nsfmax, false);
    for (int s = lane; s < nsf; s += 64) L.acc[s] = 0;
    stage<R>(m, g, w, i, e, L, seg, nullptr);
    __syncthreads();
    const R rce = (R)m.rc[e];

    // G2, symmetryFunctions.f90:220-230
    for (int f = 0; f < m.nel; ++f) {
        const int2 tb = m.g2_tab[e * kMaxEl + f];
        const int s0 = seg[f], s1 = seg[f + 1];
        if (tb.y == 0 || s1 == s0) continue;
        for (int c0 = 0; c0 < tb.y; c0 += 8) {
            const int nc = min(8, tb.y - c0);
            R acc[8];
#pragma unroll
            for (int t = 0; t < 8; ++t) acc[t] = 0;
            for (int q = s0 + lane; q < s1; q += 64) {
                const R rr = L.r[q], fcq = L.fc[q];
#pragma unroll
                for (int t = 0; t < 8; ++t)
                    if (t < nc) {
                        const int x = tb.x + c0 + t;
                        const R dr = rr - (R)m.g2_rs[x];
                        acc[t] += f_exp(-(R)m.g2_eta[x] * dr * dr) * fcq;
                    }
            }
#pragma unroll
            for (int t = 0; t < 8; ++t)
                if (t < nc) {
                    const R v = wave_sum(acc[t]);
                    if (lane == 0) L.acc[m.g2_s[tb.x + c0 + t]] += v;
                }
        }
    }
    // G4, symmetryFunctions.f90:231-305
    for (int f = 0; f < m.nel; ++f) {
        for (int h = f; h < m.nel; ++h) {
            const int2 tb = m.g4_tab[(e * kMaxEl + f) * kMaxEl + h];
            if (tb.y == 0) continue;
            const int nf = seg[f + 1] - seg[f], nh = seg[h + 1] - seg[h];
            const bool same = f == h;
            const int np = same ? nf * (nf - 1) / 2 : nf * nh;
            if (np == 0) continue;
            for (int c0 = 0; c0 < tb.y; c0 += 8) {
                const int nc = min(8, tb.y - c0);
                R acc[8];
#pragma unroll
                for (int t = 0; t < 8; ++t) acc[t] = 0;
                for (int p = lane; p < np; p += 64) {
                    int j, k;
                    decode_pair(p, same, nf, nh, seg[f], seg[h], j, k);
                    const R xj = L.dx[j], yj = L.dy[j], zj = L.dz[j];
                    const R xk = L.dx[k], yk = L.dy[k], zk = L.dz[k];
                    const R ex = xk - xj, ey = yk - yj, ez = zk - zj;   // d_jk = d_ik - d_ij
                    const R rjk2 = ex * ex + ey * ey + ez * ez;
                    const R rjk = f_sqrt(rjk2);
                    if (!(rjk < rce)) continue;
                    R fcjk, dfcjk;
                    cutoff_fn<R>(rjk, rce, fcjk, dfcjk);
                    const R rij = L.r[j], rik = L.r[k];
                    const R cs = (xj * xk + yj * yk + zj * zk) / (rij * rik);
                    const R F = L.fc[j] * L.fc[k] * fcjk;
                    const R r2 = rij * rij + rik * rik + rjk2;
                    R E = 0;
#pragma unroll
                    for (int t = 0; t < 8; ++t)
                        if (t < nc) {
                            const int x = tb.x + c0 + t;
                            if (t == 0 || m.g4_neweta[x]) E = f_exp(-(R)m.g4_eta[x] * r2);
                            const R b = R(1) + (R)m.g4_lam[x] * cs;
                            const R C = b * zpow<R>(b, m.g4_ipow[x], (R)m.g4_zeta[x]);
                            acc[t] += (R)m.g4_pre[x] * C * E * F;
                        }
                }
#pragma unroll
                for (int t = 0; t < 8; ++t)
                    if (t < nc) {
                        const R v = wave_sum(acc[t]);
                        if (lane == 0) L.acc[m.g4_s[tb.x + c0 + t]] += v;
                    }
            }
        }
    }
    __syncthreads();
    // centre and scale, hdnnps.f90:471-481
    const int so = m.sf_off[e];
    for (int s = lane; s < m.nsfmax; s += 64)
        G[(size_t)i * m.nsfmax + s] =
            s < nsf ? (R)(((double)L.acc[s] - m.sf_mean[so + s]) / m.sf_den[so + s]) : R(0);
}

template <typename R>
__global__ __launch_bounds__(64) void k_sf_grad(DevModel m, Geo g, Work w, int nmax, const R *dEdG,
                                                R *gent)
{
    extern __shared__ __align__(16) char smem[];
    __shared__ int seg[kMaxEl + 1];
    const int i = blockIdx.x;
    const int lane = lane_id();
    const int e = w.eidx[i];
    const int nsf = m.nsf[e];
    const int so = m.sf_off[e];
    Lds<R> L = carve<R>(smem, nmax, m.nsfmax, true);
    // dE/dG of the unscaled G: the scaled G is (G - mean)/den
    for (int s = lane; s < nsf; s += 64)
        L.acc[s] = (R)((double)dEdG[(size_t)i * m.nsfmax + s] / m.sf_den[so + s]);
    stage<R>(m, g, w, i, e, L, seg, gent);
    __syncthreads();
    const int nkeep = seg[kMaxEl];
    for (int q = lane; q < nkeep; q += 64) L.gx[q] = L.gy[q] = L.gz[q] = 0;
    __syncthreads();
    const R rce = (R)m.rc[e];

    // G2: g_j = d_ij * (-dfc * sum w E + fc/r * sum w 2 eta (r - Rs) E)
    for (int f = 0; f < m.nel; ++f) {
        const int2 tb = m.g2_tab[e * kMaxEl + f];
        const int s0 = seg[f], s1 = seg[f + 1];
        if (tb.y == 0 || s1 == s0) continue;
        for (int q = s0 + lane; q < s1; q += 64) {
            const R rr = L.r[q];
            R A = 0, B = 0;
            for (int t = 0; t < tb.y; ++t) {
                const int x = tb.x + t;
                const R eta = (R)m.g2_eta[x];
                const R dr = rr - (R)m.g2_rs[x];
                const R ws = L.acc[m.g2_s[x]] * f_exp(-eta * dr * dr);
                A += ws;
                B += ws * R(2) * eta * dr;
            }
            const R c = -L.dfc[q] * A + L.fc[q] * B / rr;
            L.gx[q] += c * L.dx[q];
            L.gy[q] += c * L.dy[q];
            L.gz[q] += c * L.dz[q];
        }
    }
    __syncthreads();
    // G4: per triplet, S1 = sum w P C E, S2 = sum w P C E 2 eta,
    // S3 = sum w P b^(zeta-1) lambda zeta E, then
    // g_j = S1 dF/dR_j + S2 F (d_ij - d_jk) + S3 F dcos/dR_j, same for k.
    for (int f = 0; f < m.nel; ++f) {
        for (int h = f; h < m.nel; ++h) {
            const int2 tb = m.g4_tab[(e * kMaxEl + f) * kMaxEl + h];
            if (tb.y == 0) continue;
            const int nf = seg[f + 1] - seg[f], nh = seg[h + 1] - seg[h];
            const bool same = f == h;
            const int np = same ? nf * (nf - 1) / 2 : nf * nh;
            for (int p = lane; p < np; p += 64) {
                int j, k;
                decode_pair(p, same, nf, nh, seg[f], seg[h], j, k);
                const R xj = L.dx[j], yj = L.dy[j], zj = L.dz[j];
                const R xk = L.dx[k], yk = L.dy[k], zk = L.dz[k];
                const R ex = xk - xj, ey = yk - yj, ez = zk - zj;
                const R rjk2 = ex * ex + ey * ey + ez * ez;
                const R rjk = f_sqrt(rjk2);
                if (!(rjk < rce)) continue;
                R fcjk, dfcjk;
                cutoff_fn<R>(rjk, rce, fcjk, dfcjk);
                const R rij = L.r[j], rik = L.r[k];
                const R fcij = L.fc[j], fcik = L.fc[k], dfcij = L.dfc[j], dfcik = L.dfc[k];
                const R inv = R(1) / (rij * rik);
                const R cs = (xj * xk + yj * yk + zj * zk) * inv;
                const R F = fcij * fcik * fcjk;
                const R r2 = rij * rij + rik * rik + rjk2;
                R S1 = 0, S2 = 0, S3 = 0, E = 0;
                for (int t = 0; t < tb.y; ++t) {
                    const int x = tb.x + t;
                    const R eta = (R)m.g4_eta[x];
                    if (t == 0 || m.g4_neweta[x]) E = f_exp(-eta * r2);
                    const R lam = (R)m.g4_lam[x];
                    const R b = R(1) + lam * cs;
                    const R zeta = (R)m.g4_zeta[x];
                    const R Bp = zpow<R>(b, m.g4_ipow[x], zeta);
                    const R C = Bp * b;
                    const R wpe = L.acc[m.g4_s[x]] * (R)m.g4_pre[x] * E;
                    S1 += wpe * C;
                    S2 += wpe * C * R(2) * eta;
                    S3 += wpe * Bp * lam * zeta;
                }
                const R aj = fcik * dfcjk * fcij, bj = fcik * dfcij * fcjk;   // dF/dR_j = aj d_jk - bj d_ij
                const R ak = fcij * dfcjk * fcik, bk = fcij * dfcik * fcjk;   // dF/dR_k = -(ak d_jk + bk d_ik)
                const R cj = cs / (rij * rij), ck = cs / (rik * rik);
                const R S2F = S2 * F, S3F = S3 * F;
                // g_j
                const R gjx = S1 * (aj * ex - bj * xj) + S2F * (xj - ex) + S3F * (-xk * inv + xj * cj);
                const R gjy = S1 * (aj * ey - bj * yj) + S2F * (yj - ey) + S3F * (-yk * inv + yj * cj);
                const R gjz = S1 * (aj * ez - bj * zj) + S2F * (zj - ez) + S3F * (-zk * inv + zj * cj);
                // g_k
                const R gkx = -S1 * (ak * ex + bk * xk) + S2F * (xk + ex) + S3F * (-xj * inv + xk * ck);
                const R gky = -S1 * (ak * ey + bk * yk) + S2F * (yk + ey) + S3F * (-yj * inv + yk * ck);
                const R gkz = -S1 * (ak * ez + bk * zk) + S2F * (zk + ez) + S3F * (-zj * inv + zk * ck);
                atomicAdd(&L.gx[j], gjx);
                atomicAdd(&L.gy[j], gjy);
                atomicAdd(&L.gz[j], gjz);
                atomicAdd(&L.gx[k], gkx);
                atomicAdd(&L.gy[k], gky);
                atomicAdd(&L.gz[k], gkz);
            }
        }
    }
    __syncthreads();
    // per-entry gradient rows and the atom's virial -sum_n g_n (x) d_in
    double v[9];
#pragma unroll
    for (int t = 0; t < 9; ++t) v[t] = 0;
    for (int q = lane; q < nkeep; q += 64) {
        const R gx = L.gx[q], gy = L.gy[q], gz = L.gz[q];
        const size_t en = (size_t)L.ent[q];
        gent[3 * en + 0] = gx;
        gent[3 * en + 1] = gy;
        gent[3 * en + 2] = gz;
        const double dxq = L.dx[q], dyq = L.dy[q], dzq = L.dz[q];
        v[0] -= gx * dxq; v[1] -= gx * dyq; v[2] -= gx * dzq;
        v[3] -= gy * dxq; v[4] -= gy * dyq; v[5] -= gy * dzq;
        v[6] -= gz * dxq; v[7] -= gz * dyq; v[8] -= gz * dzq;
    }
#pragma unroll
    for (int t = 0; t < 9; ++t) {
        const double s = wave_sum(v[t]);
        if (lane == 0) w.vir[9 * (size_t)i + t] = s;
    }
}

template <typename R>
void launch_sf_values(const DevModel &m, const Geo &g, Work &w, int nmax, R *G, hipStream_t s)
{
    const size_t sh = lds_bytes<R>(nmax, m.nsfmax, false);
    hipLaunchKernelGGL(k_sf_values<R>, dim3(w.nat), dim3(64), sh, s, m, g, w, nmax, G);
}

template <typename R>
void launch_sf_grad(const DevModel &m, const Geo &g, Work &w, int nmax, const R *dEdG, R *gent,
                    hipStream_t s)
{
    const size_t sh = lds_bytes<R>(nmax, m.nsfmax, true);
    hipLaunchKernelGGL(k_sf_grad<R>, dim3(w.nat), dim3(64), sh, s, m, g, w, nmax, dEdG, gent);
}

template void launch_sf_values<double>(const DevModel &, const Geo &, Work &, int, double *, hipStream_t);
template void launch_sf_values<float>(const DevModel &, const Geo &, Work &, int, float *, hipStream_t);
template void launch_sf_grad<double>(const DevModel &, const Geo &, Work &, int, const double *, double *,
                                     hipStream_t);
template void launch_sf_grad<float>(const DevModel &, const Geo &, Work &, int, const float *, float *,
                                    hipStream_t);

size_t sf_lds_bytes(bool fp64, int nmax, int nsfmax, bool grad)
{
    return fp64 ? lds_bytes<double>(nmax, nsfmax, grad) : lds_bytes<float>(nmax, nsfmax, grad);
}

}  // namespace hdnnp
