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
// within its cutoff are staged in LDS, grouped by element, as two 4-vectors
// each: (dx, dy, dz, 1/r) and (r, fc, fc'/r, r^2). Triplets whose third side
// reaches the cutoff contribute nothing (fc(r_jk) = 0), so both passes first
// compact the candidate triplets with r_jk < rc into an LDS list.
//   Pass A (values) needs one sum per function, so it runs per element-pair
//   block (f, h) of the neighbours: the block's candidates are enumerated
//   with one division each (circulant rows for f = h, diagonals for f != h),
//   and the block's SF parameters are wave-uniform, in scalar registers
//   (loaded through the constant address space, so uniform indices compile to
//   scalar loads); each lane accumulates per function, one wave reduction per
//   function and block.
//   (The FP32 pass with lambda = +-1 and zeta >= 1 is the block sweep below:
//   the blocks' triplets with r_jk < rc into a list in block order, then the
//   lanes dealt out to the blocks with the functions and sums in registers.)
//   Pass B (gradients) needs no per-function sums, so it runs over the
//   triplets of all blocks at once, in circulant order, with every lane busy
//   (after the FP32 block sweep: over its list, lanes dealt out to the
//   blocks, g4_grad_blocks); each lane reads its block's parameters, with
//   dE/dG folded in, from LDS tables (or registers), forms per triplet
//     S1 = sum w P C E, S2 = sum w P C E 2 eta, S3 = sum w P b^(zeta-1) lambda zeta E
//   and adds g_j = S1 dF/dR_j + S2 F (d_ij - d_jk) + S3 F dcos/dR_j (same for
//   k) to FP64 per-neighbour LDS accumulators.
// zeta = 2^k (the usual choice) takes k squarings; other zeta go through the
// integer power or pow. Tiles of several atoms of one element per wavefront
// are supported (AtomsPerWave) but measured slower here: LDS per atom limits
// the waves in flight.
// The FP32 path uses the hardware exp2/rcp/sqrt approximations; the FP64 path
// uses exp and tanh polynomials of about 1 ulp (fmath.hpp) and the device
// library's sqrt, pow and cospi.
#include <rocprim/intrinsics/warp_shuffle.hpp>

#include <cstdlib>
#include <cstring>

#include "device.hpp"
#include "fmath.hpp"

namespace hdnnp {

// triplet candidates compacted per LDS chunk: FP32 takes the smaller chunk
// for one more wavefront per CU; FP64, bound by its staging arrays either
// way, the larger one for fewer rounds
template <typename R> constexpr int kPairChunk = sizeof(R) == 4 ? 256 : 512;

// The compact triplet lists the passes hand over hold 16-bit entries j | k
// << 8 when the atom has at most 256 staged neighbours (both passes know n),
// else the 32-bit packing j | k << 10 | block << 20
__device__ __forceinline__ bool trip16(int n) { return n <= 256; }

// ---------------------------------------------------------------- math
template <typename R> struct M;
template <> struct M<double> {
    static __device__ double sqrt(double x) { return ::sqrt(x); }
    static __device__ double rcp(double x) { return 1.0 / x; }
    // tanh of the cutoffs, whose argument 1 - r/rc lies in [0, 1]
    static __device__ double tanh(double x) { return tanh01(x); }
    static __device__ double pow(double x, double y) { return ::pow(x, y); }
    // exp(-eta x) from the table value a = -eta
    static __device__ double expsc(double a, double x) { return exp_f64(a * x); }
    static __device__ double exp(double x) { return exp_f64(x); }
    static __device__ double cospi(double x) { return ::cospi(x); }
    static __device__ double sinpi(double x) { return ::sinpi(x); }
};
template <> struct M<float> {
    static __device__ float sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
    static __device__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
    // tanh on [0, 1] through one exponential: e = exp(-2x) = exp2(-2 log2(e) x)
    static __device__ float tanh(float x)
    {
        const float e = __builtin_amdgcn_exp2f(x * -2.8853900817779268f);
        return (1.0f - e) * __builtin_amdgcn_rcpf(1.0f + e);
    }
    static __device__ float pow(float x, float y) { return __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x)); }
    // exp(-eta x) from the table value a = -eta log2(e)
    static __device__ float expsc(float a, float x) { return __builtin_amdgcn_exp2f(a * x); }
    static __device__ float exp(float x) { return __builtin_amdgcn_exp2f(x * 1.4426950408889634f); }
    static __device__ float cospi(float x) { return ::cospif(x); }
    static __device__ float sinpi(float x) { return ::sinpif(x); }
};

// The model tables are not written while a kernel runs: reading them through
// the constant address space lets the compiler use scalar loads wherever the
// index is wave-uniform.
#define CAS __attribute__((address_space(4)))
template <typename T> __device__ __forceinline__ T cld(const T *p, int x) { return ((const CAS T *)p)[x]; }
// (start, count) of an int2 table entry
__device__ __forceinline__ void cld2(const int2 *p, int x, int &a, int &b)
{
    const CAS int *q = (const CAS int *)p;
    a = q[2 * x];
    b = q[2 * x + 1];
}
__device__ __forceinline__ ElemInfo cld_einfo(const ElemInfo *p, int e)
{
    const CAS ElemInfo *q = (const CAS ElemInfo *)p + e;
    ElemInfo r;
    r.nsf = q->nsf;
    r.sf_off = q->sf_off;
    r.g2_start = q->g2_start;
    r.g2_count = q->g2_count;
    r.g4_start = q->g4_start;
    r.g4_count = q->g4_count;
    r.rc = q->rc;
    r.atom_energy = q->atom_energy;
    return r;
}

template <typename R> __device__ __forceinline__ R tabc(const double *d, const float *f, int x);
template <> __device__ __forceinline__ double tabc<double>(const double *d, const float *, int x) { return cld(d, x); }
template <> __device__ __forceinline__ float tabc<float>(const double *, const float *f, int x) { return cld(f, x); }

// exponent scale of exp(-eta x) in the form M<R>::expsc expects
template <typename R> __device__ __forceinline__ R g2_ea(const DevModel &m, int x);
template <> __device__ __forceinline__ double g2_ea<double>(const DevModel &m, int x) { return -cld(m.g2_eta, x); }
template <> __device__ __forceinline__ float g2_ea<float>(const DevModel &m, int x) { return cld(m.g2f_ea, x); }
template <typename R> __device__ __forceinline__ R g4_ea(const DevModel &m, int x);
template <> __device__ __forceinline__ double g4_ea<double>(const DevModel &m, int x) { return -cld(m.g4_eta, x); }
template <> __device__ __forceinline__ float g4_ea<float>(const DevModel &m, int x) { return cld(m.g4f_ea, x); }

// wave-uniform copies (scalar registers)
__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ float uni(float x)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, x)));
}
__device__ __forceinline__ double uni(double x)
{
    const long long v = __builtin_bit_cast(long long, x);
    const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)v);
    const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(v >> 32));
    return __builtin_bit_cast(double, ((long long)hi << 32) | lo);
}

// Cutoff functions of RuNNer's cutoff_type (x = r/rc < 1; dfc = fc'(r)/r):
//   0 hard 1, 1 cosine (cos(pi x) + 1)/2, 2 tanh^3(1 - x) (the reference's,
//   symmetryFunctions.f90:310-327), 3 tanh^3(1 - x) / tanh^3(1), 4 exp(1 -
//   1/(1 - x^2)), 5-8 the polynomials 1 + x^2 (2x - 3), 1 + x^3 ((15 - 6x) x
//   - 10), 1 + x^4 (x (x (20x - 70) + 84) - 35), 1 + x^5 (x (x ((315 - 70x)
//   x - 540) + 420) - 126) (RuNNer's / n2p2's numbering). The type is one per
//   model (wave-uniform).
constexpr double kTanhNorm = 2.2637537952253513;   // 1 / tanh^3(1)

// The cutoff is evaluated without contracting products into FMAs: its
// rounding is then the same in every kernel instance that inlines it (each
// staging capacity), so that a structure gets the same result in any batch.
template <typename R>
__device__ __forceinline__ void cutoff_other(int ct, R r, R rrc, R &fc, R &dfc)
{
#pragma clang fp contract(off)
    const R x = r * rrc;
    R f, d;   // f(x) and df/dx
    switch (ct) {
    case 0:
        f = 1;
        d = 0;
        break;
    case 1:
        f = R(0.5) * (M<R>::cospi(x) + R(1));
        d = R(-0.5 * M_PI) * M<R>::sinpi(x);
        break;
    case 3: {
        const R t = M<R>::tanh(R(1) - x), tt = t * t;
        f = R(kTanhNorm) * (tt * t);
        d = R(-3.0 * kTanhNorm) * (tt - tt * tt);
        break;
    }
    case 4: {
        const R iu = M<R>::rcp(R(1) - x * x);
        f = M<R>::exp(R(1) - iu);
        d = f * (R(-2) * x * iu * iu);
        break;
    }
    default: {
        // the polynomials are smooth steps, f = 1 - S(x) = S(1 - x) with
        // S(u) = u^k p(u); S is evaluated at the smaller of x and 1 - x, so
        // that no large terms cancel; f'(x) = -c x^(k-1) (1 - x)^(k-1)
        const R y = R(1) - x;
        const bool lo = x < R(0.5);
        const R u = lo ? x : y, u2 = u * u;
        R S;
        if (ct == 5) {
            S = u2 * (R(3) - R(2) * u);
            d = R(-6) * x * y;
        } else if (ct == 6) {
            S = u2 * u * (R(10) + u * (R(-15) + R(6) * u));
            d = R(-30) * (x * x) * (y * y);
        } else if (ct == 7) {
            S = u2 * u2 * (R(35) + u * (R(-84) + u * (R(70) - R(20) * u)));
            d = R(-140) * (x * x * x) * (y * y * y);
        } else {
            S = u2 * u2 * u * (R(126) + u * (R(-420) + u * (R(540) + u * (R(-315) + R(70) * u))));
            const R xx = x * x, yy = y * y;
            d = R(-630) * (xx * xx) * (yy * yy);
        }
        f = lo ? R(1) - S : S;
        break;
    }
    }
    fc = f;
    dfc = d * rrc * M<R>::rcp(r);
}

// the cutoff within rc (the callers test r < rc)
// tanh^3 with its fused multiply-adds written out (explicit fma calls are
// kept as written in every instance)
__device__ __forceinline__ double fmaR(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fmaR(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

template <typename R>
__device__ __forceinline__ void cutoff_in(int ct, R r, R rrc, R &fc, R &dfc)
{
#pragma clang fp contract(off)
    if (ct == 2) {
        const R t = M<R>::tanh(fmaR(-r, rrc, R(1)));
        const R tt = t * t;
        fc = tt * t;
        dfc = -(R(3) * rrc * M<R>::rcp(r)) * fmaR(-tt, tt, tt);
    } else {
        cutoff_other<R>(ct, r, rrc, fc, dfc);
    }
}
template <typename R>
__device__ __forceinline__ R cutoff_in(int ct, R r, R rrc)
{
#pragma clang fp contract(off)
    if (ct == 2) {
        const R t = M<R>::tanh(fmaR(-r, rrc, R(1)));
        return t * t * t;
    }
    R fc, dfc;
    cutoff_other<R>(ct, r, rrc, fc, dfc);
    return fc;
}

// with the range test: fc = dfc = 0 from rc on
template <typename R>
__device__ __forceinline__ void cutoff_fn(int ct, R r, R rc, R rrc, R &fc, R &dfc)
{
    if (r < rc) {
        cutoff_in<R>(ct, r, rrc, fc, dfc);
    } else {
        fc = 0;
        dfc = 0;
    }
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

// FP32: the same steps with lanes that have no source reading 0 (the
// identity of the add), so that each step is one v_add_f32 with a DPP
// operand; the adds of lane 63 and their order are those above
template <int CTRL>
__device__ __forceinline__ float dpp_or_zero(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xf, 0xf, true));
}
template <>
__device__ __forceinline__ float wave_total<float>(float v)
{
#pragma clang fp contract(off)
    v += dpp_or_zero<0xb1>(v);
    v += dpp_or_zero<0x4e>(v);
    v += dpp_or_zero<0x124>(v);
    v += dpp_or_zero<0x128>(v);
    v += dpp_or_zero<0x142>(v);
    v += dpp_or_zero<0x143>(v);
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

__device__ __forceinline__ int mbcnt(unsigned long long m)
{
    return __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

// ---------------------------------------------------------------- LDS layout
// 32-bit pointers into LDS: they keep the LDS address space, so accesses
// compile to ds_* instructions
#define LDS __attribute__((address_space(3)))

template <typename R> struct Vec4 { typedef R T __attribute__((ext_vector_type(4))); };
typedef int I2 __attribute__((ext_vector_type(2)));
typedef int I4 __attribute__((ext_vector_type(4)));
typedef float F2 __attribute__((ext_vector_type(2)));

// Atoms per wavefront (GA, the tile): several atoms of one element can share
// a wavefront (64/GA lanes each; the SF parameters stay wave-uniform), which
// fills the lanes when atoms have few neighbours, but the LDS per atom then
// limits the waves in flight. One atom per wavefront is the default and
// measured fastest for 50-100 neighbours; 2 and 4 are built for staging
// capacities up to 128 and chosen per context (Engine::set_tile, autotune).

// One atom's LDS region.
template <typename R>
struct Lds {
    typedef typename Vec4<R>::T V;
    LDS V *a;                     // (dx, dy, dz, 1/r) of each staged neighbour
    LDS V *b;                     // (r, fc, fc'/r, r^2)
    LDS unsigned char *el;        // element rank
    LDS unsigned short *ent;      // list slot within the atom's list (< cap <= 1024)
    LDS int *pairs;               // compacted triplets: j | k << 16 (per block), j | k << 10 | block << 20
                                  // (all pairs)
    LDS int *seg;                 // [kMaxEl+1] element segments of the staged neighbours
    LDS double *gx, *gy, *gz;     // gradients: per-neighbour accumulators (FP64: ds_add_f64
                                  // is much faster than ds_add_f32 here)
    LDS R *acc;                   // values: G sums; gradients: dE/dG of the raw G
    // gradients: G4 functions of the centre element (local index u), block table
    LDS V *sp0;                   // (exp scale, lambda, q1 = w P, q2 = w P 2 eta)
    LDS R *sp1;                   // q3 = w P lambda zeta
    LDS int *smeta;               // zeta = 2^k: k, else 0xFF; | new-eta flag << 8
    LDS I2 *btab;                 // block f*kMaxEl+h (f <= h): (first u, count)
    // values over all pairs (FP32, one atom per wavefront): the function
    // sums in FP64, added to with LDS atomics; sp0 = (exp scale, lambda,
    // 2^(1-zeta), zeta - 1), smeta also holds the function's slot << 16, btab
    LDS double *accd;
};

// Everything sized by the staging capacity nmax comes first, so that with
// nmax a compile-time constant these arrays sit at constant offsets.
struct LdsOffsets {
    size_t a, b, el, ent, pairs, seg, g, acc, accd, sp0, sp1, smeta, btab, total;
};

// triplets per compacted chunk of the all-pairs values pass
constexpr int kAllPairsChunk = 512;

template <typename R>
__host__ __device__ __forceinline__ LdsOffsets lds_offsets(int nmax, int nsfmax, int g4max, int nel, bool grad,
                                                          bool allp = false)
{
    LdsOffsets o{};
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t at = off;
        off += (bytes + 15) / 16 * 16;
        return at;
    };
    o.a = take(sizeof(typename Vec4<R>::T) * (size_t)nmax);
    o.b = take(sizeof(typename Vec4<R>::T) * (size_t)nmax);
    o.el = take((size_t)nmax);
    o.ent = take(sizeof(unsigned short) * (size_t)nmax);
    o.pairs = take(sizeof(int) * (allp ? kAllPairsChunk : kPairChunk<R>));
    o.seg = take(sizeof(int) * (kMaxEl + 1));
    o.g = grad ? take(sizeof(double) * 3 * (size_t)nmax) : 0;
    o.acc = take(sizeof(R) * (size_t)nsfmax);
    if (allp && sizeof(R) == 4) o.accd = take(sizeof(double) * (size_t)nsfmax);   // FP64: acc
    if (grad || allp) {
        o.sp0 = take(sizeof(typename Vec4<R>::T) * (size_t)g4max);
        o.sp1 = grad ? take(sizeof(R) * (size_t)g4max) : 0;
        o.smeta = take(sizeof(int) * (size_t)g4max);
        o.btab = take(sizeof(I2) * (size_t)nel * (size_t)nel);   // block f*nel+h, f <= h
    }
    o.total = off;
    return o;
}

// bytes of dynamic LDS of a wavefront (GA atom regions)
template <typename R>
__host__ __device__ inline size_t lds_bytes(int nmax, int ga, int nsfmax, int g4max, int nel, bool grad,
                                            bool allp = false)
{
    return ga * lds_offsets<R>(nmax, nsfmax, g4max, nel, grad, allp).total;
}

template <typename R>
__device__ __forceinline__ Lds<R> make_lds(int nmax, int nsfmax, int g4max, int nel, bool grad, LDS char *base,
                                           bool allp = false)
{
    const LdsOffsets o = lds_offsets<R>(nmax, nsfmax, g4max, nel, grad, allp);
    Lds<R> L{};
    L.a = (LDS typename Lds<R>::V *)(base + o.a);
    L.b = (LDS typename Lds<R>::V *)(base + o.b);
    L.el = (LDS unsigned char *)(base + o.el);
    L.ent = (LDS unsigned short *)(base + o.ent);
    L.pairs = (LDS int *)(base + o.pairs);
    L.seg = (LDS int *)(base + o.seg);
    if (grad) {
        LDS double *pd = (LDS double *)(base + o.g);
        L.gx = pd;
        L.gy = pd + nmax;
        L.gz = pd + 2 * nmax;
    }
    L.acc = (LDS R *)(base + o.acc);
    if (allp) L.accd = (LDS double *)(base + (sizeof(R) == 4 ? o.accd : o.acc));
    if (grad || allp) {
        L.sp0 = (LDS typename Lds<R>::V *)(base + o.sp0);
        if (grad) L.sp1 = (LDS R *)(base + o.sp1);
        L.smeta = (LDS int *)(base + o.smeta);
        L.btab = (LDS I2 *)(base + o.btab);
    }
    return L;
}

// ---------------------------------------------------------------- lane groups
// GL lanes per atom: group = lane / GL. Ballot bits of the own group, sums
// and maxima over the group / the wavefront.
template <int GL>
__device__ __forceinline__ unsigned long long group_bits(unsigned long long mk)
{
    if (GL == 64) return mk;
    return (mk >> (lane_id() / GL * GL)) & ((1ull << (GL & 63)) - 1);
}

template <int GL, typename R>
__device__ __forceinline__ R group_sum(R v)
{
    if (GL == 64) return wave_total(v);
#pragma unroll
    for (int o = GL / 2; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// a group-uniform value: wave-uniform (scalar register) when one group fills the wavefront
template <int GL>
__device__ __forceinline__ int gu(int v) { return GL == 64 ? uni(v) : v; }

__device__ __forceinline__ int wave_max(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
    return uni(v);
}

// GA-atom tile `blk` of the element-sorted atom order: element e, first
// position t0 in perm, tn atoms (e = -1 past the last tile). One atom per
// wavefront keeps the index order (t0 = atom, no perm lookup: tn = -1).
__device__ __forceinline__ void elem_tile(const Work &w, int nel, int blk, int ga, int &e, int &t0, int &tn)
{
    if (ga == 1) {
        e = blk < w.nat ? uni(w.eidx[blk]) : -1;
        t0 = blk;
        tn = -1;
        return;
    }
    e = -1;
    t0 = tn = 0;
    int acc = 0;
    for (int f = 0; f < nel; ++f) {
        const int cnt = uni(w.elem_count[f]);
        const int nt = (cnt + ga - 1) / ga;
        if (blk < acc + nt) {
            e = f;
            t0 = uni(w.elem_off[f]) + (blk - acc) * ga;
            tn = min(ga, cnt - (blk - acc) * ga);
            return;
        }
        acc += nt;
    }
}

// Lattice (list frame) of the structure of atom i.
struct Lat9 {
    double v[9];
};
__device__ __forceinline__ Lat9 atom_lattice(const Geo &g, const Work &w, int i)
{
    Lat9 L;
    if (i < 0) {
#pragma unroll
        for (int k = 0; k < 9; ++k) L.v[k] = 0.0;
        return L;
    }
    const CAS StructGeo *G = (const CAS StructGeo *)g.sg + w.sid[i];
#pragma unroll
    for (int k = 0; k < 9; ++k) L.v[k] = G->lat[k];
    return L;
}

// Geometry of one list entry: d = R_i - (R_j + L.h) (FP64), r = |d| in
// the precision R of the symmetry functions (FP32: from the rounded d).
template <typename R = double>
__device__ __forceinline__ void entry_geometry(const Lat9 &lt, const Work &w, const double4 &pi, int e,
                                               double &d0, double &d1, double &d2, double &r, int &j)
{
    const double *lat = lt.v;
    j = w.ent_j[e];
    int sx, sy, sz;
    unpack_shift(w.ent_shift[e], sx, sy, sz);
    const double l0 = ((lat[6] * sz) + lat[3] * sy) + lat[0] * sx;
    const double l1 = ((lat[7] * sz) + lat[4] * sy) + lat[1] * sx;
    const double l2 = ((lat[8] * sz) + lat[5] * sy) + lat[2] * sx;
    const double4 pj = w.wpos[j];
    d0 = pi.x - (pj.x + l0);
    d1 = pi.y - (pj.y + l1);
    d2 = pi.z - (pj.z + l2);
    if constexpr (sizeof(R) == 4) {
        const float x = (float)d0, y = (float)d1, z = (float)d2;
        r = (double)M<float>::sqrt(x * x + y * y + z * z);
    } else {
        r = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
    }
}

// Stage the neighbours of atom i (i < 0: idle group) with r < rc(e) into its
// LDS region, grouped by element (seg[f]..seg[f+1]), list order within an
// element. Entries beyond the cutoff contribute nothing (fc = dfc = 0
// there); their gradient rows are zeroed. nwave: most entries of an atom of
// the wavefront.
template <typename R, int GL, int NMAX, bool KEEP>
__device__ __forceinline__ void stage(const DevModel &m, const Geo &g, const Work &w, int i, int nwave, double rce,
                                      const Lds<R> &L, R *gent)
{
    // KEEP: up to 4 chunks of GL entries the counting pass keeps its geometry
    // in registers for the placing pass (pays in pass B and in the FP32
    // all-pairs pass A; elsewhere in pass A the extra registers cost more
    // occupancy than the recomputation)
    constexpr int NC = (NMAX + GL - 1) / GL;
    constexpr bool kKeep = KEEP && NC <= 4;
    constexpr int NK = kKeep ? NC : 1;
    const int gl = lane_id() % GL;
    const int n = i >= 0 ? w.count[i] : 0;
    const int off = i >= 0 ? i * w.cap : 0;
    const double4 pi = i >= 0 ? w.wpos[i] : make_double4(0, 0, 0, 0);
    const Lat9 lt = atom_lattice(g, w, i);
    const R rrc = R(1) / (R)rce;
    double k0[NK], k1[NK], k2[NK], kr[NK];
    int kf[NK];
    int cnt[kMaxEl];
#pragma unroll
    for (int f = 0; f < kMaxEl; ++f) cnt[f] = 0;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        const int c0 = c * GL;
        if (c0 >= nwave) break;
        const int q = c0 + gl;
        int f = -1;
        double d0 = 0, d1 = 0, d2 = 0, r = 1e300;
        if (q < n) {
            int j;
            entry_geometry<R>(lt, w, pi, off + q, d0, d1, d2, r, j);
            f = r < rce ? w.eidx[j] : -1;
        }
        if (kKeep) {
            k0[kKeep ? c : 0] = d0;
            k1[kKeep ? c : 0] = d1;
            k2[kKeep ? c : 0] = d2;
            kr[kKeep ? c : 0] = r;
            kf[kKeep ? c : 0] = f;
        }
#pragma unroll
        for (int ff = 0; ff < kMaxEl; ++ff)
            if (ff < m.nel) cnt[ff] += __popcll(group_bits<GL>(__ballot(f == ff)));
    }
    int start[kMaxEl], run[kMaxEl];
    int acc = 0;
#pragma unroll
    for (int f = 0; f < kMaxEl; ++f) {
        start[f] = acc;
        run[f] = 0;
        if (gl == 0) L.seg[f] = acc;
        acc += (f < m.nel) ? cnt[f] : 0;
    }
    if (gl == 0) L.seg[kMaxEl] = acc;
    const unsigned long long below = (1ull << gl) - 1;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        const int c0 = c * GL;
        if (c0 >= nwave) break;
        const int q = c0 + gl;
        int f;
        double d0, d1, d2, r;
        if (kKeep) {
            d0 = k0[kKeep ? c : 0];
            d1 = k1[kKeep ? c : 0];
            d2 = k2[kKeep ? c : 0];
            r = kr[kKeep ? c : 0];
            f = kf[kKeep ? c : 0];
        } else {
            f = -1;
            d0 = d1 = d2 = 0;
            r = 1e300;
            if (q < n) {
                int j;
                entry_geometry<R>(lt, w, pi, off + q, d0, d1, d2, r, j);
                f = r < rce ? w.eidx[j] : -1;
            }
        }
        int pos = -1;
#pragma unroll
        for (int ff = 0; ff < kMaxEl; ++ff) {
            if (ff < m.nel) {
                const unsigned long long mk = group_bits<GL>(__ballot(f == ff));
                if (f == ff) pos = start[ff] + run[ff] + __popcll(mk & below);
                run[ff] += __popcll(mk);
            }
        }
        if (f >= 0) {
            const R rr = (R)r;
            R fc, dfc;
            cutoff_fn<R>(m.cutoff_type, rr, (R)rce, rrc, fc, dfc);
            typename Lds<R>::V va, vb;
            va.x = (R)d0;
            va.y = (R)d1;
            va.z = (R)d2;
            va.w = sizeof(R) == 4 ? M<R>::rcp(rr) : (R)(1.0 / r);
            vb.x = rr;
            vb.y = fc;
            vb.z = dfc;
            vb.w = rr * rr;
            L.a[pos] = va;
            L.b[pos] = vb;
            L.ent[pos] = (unsigned short)q;
            L.el[pos] = (unsigned char)f;
        } else if (q < n && gent) {
            gent[3 * (size_t)(off + q) + 0] = 0;
            gent[3 * (size_t)(off + q) + 1] = 0;
            gent[3 * (size_t)(off + q) + 2] = 0;
        }
    }
}

// ---------------------------------------------------------------- G4
// b^n, n >= 0 wave-uniform: square-and-multiply with scalar control (the
// integer branch of the reference's power, symmetryFunctions.f90:265-273)
template <typename R>
__device__ __forceinline__ R upow(R b, int n)
{
    R r = R(1);
    while (true) {
        if (n & 1) r *= b;
        n >>= 1;
        if (!n) break;
        b *= b;
    }
    return r;
}

// b^(zeta-1): ip = zeta-1 for integral zeta, else kNonIntegralZeta (wave-uniform)
template <typename R>
__device__ __forceinline__ R angpow(R b, int ip, R zm1)
{
    if (ip >= 0) return upow(b, ip);
    if (ip == kNonIntegralZeta) return b > R(0) ? M<R>::pow(b, zm1) : (zm1 > R(0) ? R(0) : R(1));
    return R(1) / upow(b, -ip);
}

// zeta = 2^sq (sq wave-uniform): bz = b^zeta by sq squarings and
// bm = b^(zeta-1) = prod_{i<sq} b^(2^i), no selects
template <typename R>
__device__ __forceinline__ void pow2z(R b, int sq, R &bz, R &bm)
{
    R r = R(1);
    for (int k = 0; k < sq; ++k) {
        r *= b;
        b *= b;
    }
    bz = b;
    bm = r;
}
template <typename R>
__device__ __forceinline__ R pow2z(R b, int sq)
{
    for (int k = 0; k < sq; ++k) b *= b;
    return b;
}

// floor(p / n) of a candidate index, for n <= 1024 and p < 1024 n: the float
// quotient (p + 1/2) * (1/n) errs by less than 2e-7 (p/n) <= 2e-4, while
// (p + 1/2)/n lies at least 1/(2n) >= 4.9e-4 from an integer, so truncation
// gives the row exactly (checked for every such n and p, also with 1/n off
// by one ulp); p - q n then takes the full-rate 24-bit multiply
__device__ __forceinline__ int row_of(int p, float inv_n) { return (int)(((float)p + 0.5f) * inv_n); }

// The candidate triplets [cb, ce) of one element-pair block of each atom,
// compacted to those with r_jk < rc (the others give fc(r_jk) = 0). Same
// element (segment sa, na): circulant rows, p = q*na + a is
// (a, a + q + 1 mod na) (for even na the last row holds (a, a + na/2),
// a < na/2). Two elements (na <= nb): diagonals, p = q*na + a is
// (a, a + q mod nb). One division by na, through its reciprocal. cew: the
// largest ce of the wavefront. Returns the atom's count.
template <typename R, int GL>
__device__ __forceinline__ int compact(const Lds<R> &L, int cb, int ce, int cew, bool same, int sa, int na,
                                       float inv_na, int sb, int nb, R rc2)
{
    const int gl = lane_id() % GL;
    const unsigned long long below = (1ull << gl) - 1;
    int nv = 0;
    for (int p0 = cb; p0 < cew; p0 += GL) {
        const int p = p0 + gl;
        bool ok = false;
        int pk = 0;
        if (p < ce) {
            const int q = row_of(p, inv_na);
            const int a = p - __mul24(q, na);
            int j, k;
            if (same) {
                int b = a + q + 1;
                if (b >= na) b -= na;
                j = sa + a;
                k = sa + b;
            } else {
                int b = a + q;
                if (b >= nb) b -= nb;
                j = sa + a;
                k = sb + b;
            }
            const typename Lds<R>::V vj = L.a[j], vk = L.a[k];
            const R ex = vk.x - vj.x, ey = vk.y - vj.y, ez = vk.z - vj.z;
            ok = ex * ex + ey * ey + ez * ez < rc2;
            pk = j | (k << 16);
        }
        const unsigned long long mk = group_bits<GL>(__ballot(ok));
        if (ok) L.pairs[nv + __popcll(mk & below)] = pk;
        nv += __popcll(mk);
    }
    return nv;
}

constexpr int kSfChunk = 4;        // G4 functions per pass over a compacted list

// Parameters of up to kSfChunk G4 functions, wave-uniform.
template <typename R>
struct G4Chunk {
    R ea[kSfChunk], lam[kSfChunk], zm1[kSfChunk], pre[kSfChunk];
    int ip[kSfChunk], nw[kSfChunk], s[kSfChunk], sq[kSfChunk];
};

template <typename R, int NC>
__device__ __forceinline__ void load_chunk(const DevModel &m, int x0, G4Chunk<R> &c)
{
#pragma unroll
    for (int t = 0; t < NC; ++t) {
        const int x = x0 + t;
        c.ea[t] = uni(g4_ea<R>(m, x));
        c.lam[t] = uni(tabc<R>(m.g4_lam, m.g4f_lam, x));
        c.zm1[t] = uni(tabc<R>(m.g4_zeta, m.g4f_zeta, x)) - R(1);
        c.pre[t] = uni(tabc<R>(m.g4_pre, m.g4f_pre, x));
        c.ip[t] = uni(cld(m.g4_ipow, x));
        c.sq[t] = uni(cld(m.g4_sq, x));
        c.nw[t] = t == 0 ? 1 : uni(cld(m.g4_neweta, x));
        c.s[t] = uni(cld(m.g4_s, x));
    }
}

// Pass A over the compacted lists for NC functions of a block.
template <typename R, int GL, int NC, bool FAST>
__device__ __forceinline__ void g4_values(const DevModel &m, const Lds<R> &L, int nv, int nvw, int x0, R rrc)
{
    G4Chunk<R> c;
    load_chunk<R, NC>(m, x0, c);
    const int gl = lane_id() % GL;
    R acc[NC];
#pragma unroll
    for (int t = 0; t < NC; ++t) acc[t] = 0;
    for (int q0 = 0; q0 < nvw; q0 += GL) {
        const int q = q0 + gl;
        if (q >= nv) continue;
        const int pk = L.pairs[q];
        const int j = pk & 0xFFFF, k = pk >> 16;
        const typename Lds<R>::V aj = L.a[j], ak = L.a[k], bj = L.b[j], bk = L.b[k];
        const R ex = ak.x - aj.x, ey = ak.y - aj.y, ez = ak.z - aj.z;   // d_jk = d_ik - d_ij
        const R rjk2 = ex * ex + ey * ey + ez * ez;
        const R fcjk = cutoff_in<R>(m.cutoff_type, M<R>::sqrt(rjk2), rrc);
        const R cs = (aj.x * ak.x + aj.y * ak.y + aj.z * ak.z) * aj.w * ak.w;
        const R F = bj.y * bk.y * fcjk;
        const R r2 = bj.w + bk.w + rjk2;
        R EF = 0;
#pragma unroll
        for (int t = 0; t < NC; ++t) {
            if (c.nw[t]) EF = M<R>::expsc(c.ea[t], r2) * F;
            const R b = R(1) + c.lam[t] * cs;
            acc[t] += (FAST ? pow2z<R>(b, c.sq[t]) : b * angpow<R>(b, c.ip[t], c.zm1[t])) * EF;
        }
    }
#pragma unroll
    for (int t = 0; t < NC; ++t) {
        const R v = group_sum<GL>(acc[t]);
        if (gl == 0) L.acc[c.s[t]] += c.pre[t] * v;
    }
}

// Pass B over the compacted lists for NC functions of a block: S1..S3 of the
// triplet, then g_j = cjj d_ij + cjk d_ik and g_k = ckk d_ik + cjk d_ij,
// added to the FP64 per-neighbour accumulators.
template <typename R, int GL, int NC, bool FAST>
__device__ __forceinline__ void g4_grad(const DevModel &m, const Lds<R> &L, int nv, int nvw, int x0, R rrc)
{
    G4Chunk<R> c;
    load_chunk<R, NC>(m, x0, c);
    R q1[NC], q2[NC], q3[NC];     // w P, w P 2 eta, w P lambda zeta (w = dE/dG of the raw G)
#pragma unroll
    for (int t = 0; t < NC; ++t) {
        q1[t] = L.acc[c.s[t]] * c.pre[t];
        q2[t] = q1[t] * R(2) * uni(tabc<R>(m.g4_eta, m.g4f_eta, x0 + t));
        q3[t] = q1[t] * c.lam[t] * (c.zm1[t] + R(1));
    }
    const int gl = lane_id() % GL;
    for (int q0 = 0; q0 < nvw; q0 += GL) {
        const int q = q0 + gl;
        if (q >= nv) continue;
        const int pk = L.pairs[q];
        const int j = pk & 0xFFFF, k = pk >> 16;
        const typename Lds<R>::V aj = L.a[j], ak = L.a[k], bj = L.b[j], bk = L.b[k];
        const R ex = ak.x - aj.x, ey = ak.y - aj.y, ez = ak.z - aj.z;   // d_jk
        const R rjk2 = ex * ex + ey * ey + ez * ez;
        R fcjk, dfcjk;
        cutoff_in<R>(m.cutoff_type, M<R>::sqrt(rjk2), rrc, fcjk, dfcjk);
        const R irj = aj.w, irk = ak.w;
        const R fcij = bj.y, fcik = bk.y, dfcij = bj.z, dfcik = bk.z;
        const R inv = irj * irk;
        const R cs = (aj.x * ak.x + aj.y * ak.y + aj.z * ak.z) * inv;
        const R F = fcij * fcik * fcjk;
        const R r2 = bj.w + bk.w + rjk2;
        R S1 = 0, S2 = 0, S3 = 0, E = 0;
#pragma unroll
        for (int t = 0; t < NC; ++t) {
            if (c.nw[t]) E = M<R>::expsc(c.ea[t], r2);
            const R b = R(1) + c.lam[t] * cs;
            R bz, bm;
            if (FAST) {
                pow2z<R>(b, c.sq[t], bz, bm);
            } else {
                bm = angpow<R>(b, c.ip[t], c.zm1[t]);
                bz = bm * b;
            }
            const R EB = E * bm;
            const R C = E * bz;
            S1 += C * q1[t];
            S2 += C * q2[t];
            S3 += EB * q3[t];
        }
        // dF/dR_j = a d_jk - bj d_ij, dF/dR_k = -(a d_jk + bk d_ik)
        const R a = fcik * dfcjk * fcij;
        const R bjj = fcik * dfcij * fcjk, bkk = fcij * dfcik * fcjk;
        const R S2F = S2 * F, S3F = S3 * F;
        const R beta = S1 * a - S2F;
        const R cjk = beta - S3F * inv;
        const R cjj = S2F - S1 * bjj + S3F * cs * irj * irj - beta;
        const R ckk = S2F - S1 * bkk + S3F * cs * irk * irk - beta;
        __hip_atomic_fetch_add(&L.gx[j], (double)(cjj * aj.x + cjk * ak.x), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        __hip_atomic_fetch_add(&L.gy[j], (double)(cjj * aj.y + cjk * ak.y), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        __hip_atomic_fetch_add(&L.gz[j], (double)(cjj * aj.z + cjk * ak.z), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        __hip_atomic_fetch_add(&L.gx[k], (double)(ckk * ak.x + cjk * aj.x), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        __hip_atomic_fetch_add(&L.gy[k], (double)(ckk * ak.y + cjk * aj.y), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        __hip_atomic_fetch_add(&L.gz[k], (double)(ckk * ak.z + cjk * aj.z), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
}

// All G4 blocks of the centre element e, for every atom of the wavefront:
// per block, the candidates in chunks of kPairChunk<R> are compacted, then the
// block's functions are evaluated kSfChunk at a time. GRAD selects pass B.
// Pass A with one atom per wavefront also writes the compacted triplets to
// tl (at most tcap) for pass B; it returns how many there are.
template <typename R, int GL, bool GRAD>
__device__ __forceinline__ int g4_blocks(const DevModel &m, const Lds<R> &L, int e, R rce, int *tl = nullptr,
                                         int tcap = 0)
{
    int tbase = 0;
    const R rce2 = rce * rce, rrc = R(1) / rce;
    for (int f = 0; f < m.nel; ++f) {
        for (int h = f; h < m.nel; ++h) {
            int tbx, tby;
            cld2(m.g4_tab, (e * kMaxEl + f) * kMaxEl + h, tbx, tby);
            tbx = uni(tbx);
            tby = uni(tby);
            if (tby == 0) continue;
            int sa = gu<GL>(L.seg[f]), na = gu<GL>(L.seg[f + 1]) - sa;
            int sb = gu<GL>(L.seg[h]), nb = gu<GL>(L.seg[h + 1]) - sb;
            const bool same = f == h;
            if (!same && na > nb) {
                int t = sa; sa = sb; sb = t;
                t = na; na = nb; nb = t;
            }
            const int np = same ? na * (na - 1) / 2 : na * nb;
            const int npw = GL == 64 ? np : wave_max(np);
            if (npw == 0) continue;
            const float inv_na = 1.0f / (float)max(na, 1);
            bool fast = true;    // every zeta a power of two
            for (int t = 0; t < tby; ++t) fast = fast && uni(cld(m.g4_sq, tbx + t)) >= 0;
            for (int cb = 0; cb < npw; cb += kPairChunk<R>) {
                const int nv = compact<R, GL>(L, cb, min(np, cb + kPairChunk<R>), min(npw, cb + kPairChunk<R>), same, sa,
                                              na, inv_na, sb, nb, rce2);
                const int nvw = GL == 64 ? uni(nv) : wave_max(nv);
                __syncthreads();
                if (!GRAD && GL == 64 && sizeof(R) == 8 && tl) {   // FP64 only (the engine passes none in FP32)
                    const int blk = f * m.nel + h;
                    const bool t16 = trip16(uni(L.seg[kMaxEl]));
                    for (int q = lane_id(); q < nvw; q += 64) {
                        const int pk = L.pairs[q];
                        if (tbase + q < tcap) {
                            if (t16) ((unsigned short *)tl)[tbase + q] = (unsigned short)((pk & 0xFF) | ((pk >> 16) << 8));
                            else tl[tbase + q] = (pk & 0xFFFF) | ((pk >> 16) << 10) | (blk << 20);
                        }
                    }
                    tbase += nvw;
                }
                for (int c0 = 0; c0 < tby; c0 += kSfChunk) {
                    const int x0 = tbx + c0;
#define HDNNP_G4(NC)                                                                       \
    case NC:                                                                               \
        if (GRAD) {                                                                        \
            if (fast) g4_grad<R, GL, NC, true>(m, L, nv, nvw, x0, rrc);                    \
            else g4_grad<R, GL, NC, false>(m, L, nv, nvw, x0, rrc);                        \
        } else {                                                                           \
            if (fast) g4_values<R, GL, NC, true>(m, L, nv, nvw, x0, rrc);                  \
            else g4_values<R, GL, NC, false>(m, L, nv, nvw, x0, rrc);                      \
        }                                                                                  \
        break;
                    switch (min(kSfChunk, tby - c0)) { HDNNP_G4(1) HDNNP_G4(2) HDNNP_G4(3) HDNNP_G4(4) }
#undef HDNNP_G4
                }
                __syncthreads();
            }
        }
    }
    return tbase;
}

// ---------------------------------------------------------------- pass B, all pairs
// Pass B does not need per-function sums, so it runs over the triplets of all
// blocks at once with every lane busy: the unordered pairs of the n staged
// neighbours in circulant order (row d = 1..(n-1)/2 holds (j, j+d mod n) for
// every j; for even n the row d = n/2 holds (j, j+n/2), j < n/2), so the
// lanes of one round touch distinct j and distinct k. Each lane reads the G4
// parameters of its triplet's block from the LDS tables.
__device__ __forceinline__ void circulant_pair(int p, int n, float inv_n, int &j, int &k)
{
    const int rows = (n - 1) / 2;
    int d, jj;
    if (p < rows * n) {
        const int q = row_of(p, inv_n);
        d = q + 1;
        jj = p - __mul24(q, n);
    } else {
        d = n / 2;
        jj = p - rows * n;
    }
    j = jj;
    k = jj + d;
    if (k >= n) k -= n;
}

// block f*nel+h (f <= h) of the centre element e: (first function - g4s, count)
template <typename R>
__device__ __forceinline__ void block_table(const DevModel &m, const Lds<R> &L, int e, int g4s)
{
    for (int t = lane_id(); t < m.nel * m.nel; t += 64) {
        const int f = t / m.nel, h = t - f * m.nel;
        int sx = 0, cnt = 0;
        if (f <= h) {
            cld2(m.g4_tab, (e * kMaxEl + f) * kMaxEl + h, sx, cnt);
            sx -= g4s;
        }
        L.btab[t] = I2{sx, cnt};
    }
}

// G4 tables of the centre element for the all-pairs pass, with the weights
// w = dE/dG of the raw G (in acc) folded in: sp0 = (exp scale, lambda, q1 =
// w P, q2 = w P 2 eta), sp1 = q3 = w P lambda zeta, smeta = sq | new-eta << 8,
// or (logf) the bits of zeta - 1
template <typename R>
__device__ __forceinline__ void grad_tables(const DevModel &m, const Lds<R> &L, int e, int g4s, int g4n,
                                            bool logf = false)
{
    const int lane = lane_id();
    for (int u = lane; u < g4n; u += 64) {
        const int x = g4s + u;
        const R lam = tabc<R>(m.g4_lam, m.g4f_lam, x), zeta = tabc<R>(m.g4_zeta, m.g4f_zeta, x);
        const R q1 = L.acc[cld(m.g4_s, x)] * tabc<R>(m.g4_pre, m.g4f_pre, x);
        typename Lds<R>::V p;
        p.x = g4_ea<R>(m, x);
        p.y = lam;
        p.z = q1;
        p.w = q1 * R(2) * tabc<R>(m.g4_eta, m.g4f_eta, x);
        L.sp0[u] = p;
        L.sp1[u] = q1 * lam * zeta;
        const int sq = cld(m.g4_sq, x);
        L.smeta[u] = logf ? __float_as_int((float)(zeta - R(1))) : (sq >= 0 ? sq : 0xFF) | (cld(m.g4_neweta, x) ? 1 << 8 : 0);
    }
    block_table<R>(m, L, e, g4s);
}

// G4 tables of the centre element for the all-pairs values pass: sp0 =
// (exp scale, lambda, 2^(1-zeta), zeta - 1), smeta = sq | new-eta << 8 |
// the function's slot in the atom's G << 16
template <typename R>
__device__ __forceinline__ void values_tables(const DevModel &m, const Lds<R> &L, int e, int g4s, int g4n)
{
    const int lane = lane_id();
    for (int u = lane; u < g4n; u += 64) {
        const int x = g4s + u;
        typename Lds<R>::V p;
        p.x = g4_ea<R>(m, x);
        p.y = tabc<R>(m.g4_lam, m.g4f_lam, x);
        p.z = tabc<R>(m.g4_pre, m.g4f_pre, x);
        p.w = tabc<R>(m.g4_zeta, m.g4f_zeta, x) - R(1);
        L.sp0[u] = p;
        const int sq = cld(m.g4_sq, x);
        L.smeta[u] = (sq >= 0 ? sq : 0xFF) | (cld(m.g4_neweta, x) ? 1 << 8 : 0) | (cld(m.g4_s, x) << 16);
    }
    block_table<R>(m, L, e, g4s);
}

// candidates [cb, ce) of the circulant enumeration, compacted to the
// triplets whose block has G4 functions and r_jk < rc
template <typename R>
__device__ __forceinline__ int compact_circ(const Lds<R> &L, int cb, int ce, int n, float inv_n, R rc2, int nel,
                                           int out = 0)
{
    const int lane = lane_id();
    int nv = out;
    for (int p0 = cb; p0 < ce; p0 += 64) {
        const int p = p0 + lane;
        bool ok = false;
        int pk = 0;
        if (p < ce) {
            int j, k;
            circulant_pair(p, n, inv_n, j, k);
            const int ej = L.el[j], ek = L.el[k];
            const int blk = min(ej, ek) * nel + max(ej, ek);
            const typename Lds<R>::V vj = L.a[j], vk = L.a[k];
            const R ex = vk.x - vj.x, ey = vk.y - vj.y, ez = vk.z - vj.z;
            ok = L.btab[blk].y > 0 && ex * ex + ey * ey + ez * ez < rc2;
            pk = j | (k << 10) | (blk << 20);
        }
        const unsigned long long mk = __ballot(ok);
        if (ok) L.pairs[nv + mbcnt(mk)] = pk;
        nv += __popcll(mk);
    }
    return nv - out;
}

// candidates from p0 on, compacted as compact_circ, until the next round
// could overflow cap (p0 advances past what was taken)
template <typename R>
__device__ __forceinline__ int compact_circ_fill(const Lds<R> &L, int &p0, int np, int cap, int n, float inv_n, R rc2,
                                                 int nel)
{
    int nv = 0;
    for (; p0 < np && nv <= cap - 64; p0 += 64) nv += compact_circ<R>(L, p0, min(np, p0 + 64), n, inv_n, rc2, nel, nv);
    return nv;
}

// Pass A over a compacted chunk of the all-pairs enumeration: each lane
// evaluates its triplet for the functions of the triplet's block and adds
// 2^(1-zeta) (1 + lambda cos)^zeta exp(-eta sum r^2) fc fc fc to the FP64
// function sums with LDS atomics, whose order within a wavefront is fixed
template <typename R, int CT>
__device__ __forceinline__ void g4_values_list(const DevModel &m, const Lds<R> &L, int nv, int g4s, R rrc)
{
    const int ct = CT >= 0 ? CT : m.cutoff_type;
    const int lane = lane_id();
    for (int q = lane; q < nv; q += 64) {
        const int pk = L.pairs[q];
        const int j = pk & 0x3FF, k = (pk >> 10) & 0x3FF;
        const I2 bt = L.btab[pk >> 20];
        const typename Lds<R>::V aj = L.a[j], ak = L.a[k], bj = L.b[j], bk = L.b[k];
        const R ex = ak.x - aj.x, ey = ak.y - aj.y, ez = ak.z - aj.z;   // d_jk
        const R rjk2 = ex * ex + ey * ey + ez * ez;
        const R fcjk = cutoff_in<R>(ct, M<R>::sqrt(rjk2), rrc);
        const R cs = (aj.x * ak.x + aj.y * ak.y + aj.z * ak.z) * aj.w * ak.w;
        const R F = bj.y * bk.y * fcjk;
        const R r2 = bj.w + bk.w + rjk2;
        R EF = 0;
        for (int t = 0; t < bt.y; ++t) {
            const int u = bt.x + t;
            const typename Lds<R>::V p = L.sp0[u];
            const int mt = L.smeta[u];
            if (t == 0 || (mt & (1 << 8))) EF = M<R>::expsc(p.x, r2) * F;
            const R b = R(1) + p.y * cs;
            const int sq = mt & 0xFF;
            const R P = sq != 0xFF ? pow2z<R>(b, sq) : b * angpow<R>(b, cld(m.g4_ipow, g4s + u), p.w);
            __hip_atomic_fetch_add(&L.accd[mt >> 16], (double)(p.z * (P * EF)), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
        }
    }
}

__device__ __forceinline__ float log2_base(float b);

// 1 + cos and 1 - cos of the angle between d_ij and d_ik (a = (d, 1/r)) as
// |u_j + u_k|^2 / 2 and |u_j - u_k|^2 / 2 of the unit vectors: accurate to a
// few ulp also where the angle is near 0 or pi, where 1 -+ cos in FP32 loses
// the digits that (1 -+ cos)^zeta of large zeta needs
__device__ __forceinline__ void one_pm_cos(const Vec4<float>::T &aj, const Vec4<float>::T &ak, float &bp, float &bm)
{
#pragma clang fp contract(off)
    const float ujx = aj.x * aj.w, ujy = aj.y * aj.w, ujz = aj.z * aj.w;
    const float ukx = ak.x * ak.w, uky = ak.y * ak.w, ukz = ak.z * ak.w;
    const float sx = ujx + ukx, sy = ujy + uky, sz = ujz + ukz;
    const float dx = ujx - ukx, dy = ujy - uky, dz = ujz - ukz;
    bp = 0.5f * fmaR(sz, sz, fmaR(sy, sy, sx * sx));
    bm = 0.5f * fmaR(dz, dz, fmaR(dy, dy, dx * dx));
}

// G4 functions of one element-pair block kept in registers by the passes
// that deal the lanes out to the blocks
constexpr int kRowFns = 4;
constexpr int kSumRun = 4;   // terms summed in FP32 before they go to the FP64 sums
constexpr int kFlatLen = 8;  // columns per item of the values pass's flat schedule

// ceil(c / T) for 0 <= c < 2^22 and T >= 1, with rT about 1 / T (the
// quotient of the product is off by at most one, which the remainder fixes)
__device__ __forceinline__ int ceil_div_f(int c, int T, float rT)
{
    int qq = (int)((float)c * rT);
    int rr = c - qq * T;
    if (rr >= T) {
        ++qq;
        rr -= T;
    }
    if (rr < 0) {
        --qq;
        rr += T;
    }
    return qq + (rr > 0);
}

// sum over the wavefront of v in [0, 128): one ballot per bit
__device__ __forceinline__ int wave_sum7(int v)
{
    int sum = 0;
#pragma unroll
    for (int bit = 0; bit < 7; ++bit) sum += __popcll(__ballot((v >> bit) & 1)) << bit;
    return sum;
}

// inclusive prefix sum over the lanes of the wavefront
__device__ __forceinline__ int wave_scan_incl(int v)
{
    const int lane = lane_id();
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(v, o);
        if (lane >= o) v += y;
    }
    return v;
}

// Steps T of a wavefront dealt out to nw (< 64) blocks of a list of V
// entries, block x (this lane's, x = lane < nw) holding cx: each block gets
// nlx = ceil(cx / T) lanes, and T is the fewest steps (from ceil(V / 64),
// trying two more, else ceil(V / (64 - nw)), which always fits) with at most
// 64 lanes in all
__device__ __forceinline__ void deal_lanes(int V, int nw, int cx, int &T, int &nlx)
{
    T = max((V + 63) >> 6, 1);
    nlx = ceil_div_f(cx, T, __builtin_amdgcn_rcpf((float)T));
    for (int tr = 0; tr < 2 && wave_sum7(nlx) > 64; ++tr) {
        ++T;
        nlx = ceil_div_f(cx, T, __builtin_amdgcn_rcpf((float)T));
    }
    if (wave_sum7(nlx) > 64) {   // at most V / T + nw <= 64 lanes
        T = ceil_div_f(V, 64 - nw, __builtin_amdgcn_rcpf((float)(64 - nw)));
        nlx = ceil_div_f(cx, T, __builtin_amdgcn_rcpf((float)T));
    }
}

// One triplet (j, k) of pass B: its geometry, the weighted function sums
// S1 = sum w P b^zeta E, S2 = sum w P 2 eta b^zeta E, S3 = sum w P lambda
// zeta b^(zeta-1) E by sums(aj, ak, cs, r2, S1, S2, S3), and the gradient
// rows of j and k (FP64 LDS sums)
template <typename R, int CT, typename Sums>
__device__ __forceinline__ void g4_grad_triplet(const DevModel &m, const Lds<R> &L, int j, int k, R rrc, Sums sums)
{
    const int ct = CT >= 0 ? CT : m.cutoff_type;
    const typename Lds<R>::V aj = L.a[j], ak = L.a[k], bj = L.b[j], bk = L.b[k];
    const R ex = ak.x - aj.x, ey = ak.y - aj.y, ez = ak.z - aj.z;   // d_jk
    const R rjk2 = ex * ex + ey * ey + ez * ez;
    R fcjk, dfcjk;
    cutoff_in<R>(ct, M<R>::sqrt(rjk2), rrc, fcjk, dfcjk);
    const R irj = aj.w, irk = ak.w;
    const R fcij = bj.y, fcik = bk.y, dfcij = bj.z, dfcik = bk.z;
    const R inv = irj * irk;
    const R cs = (aj.x * ak.x + aj.y * ak.y + aj.z * ak.z) * inv;
    const R F = fcij * fcik * fcjk;
    const R r2 = bj.w + bk.w + rjk2;
    R S1 = 0, S2 = 0, S3 = 0;
    sums(aj, ak, cs, r2, S1, S2, S3);
    // dF/dR_j = a d_jk - bj d_ij, dF/dR_k = -(a d_jk + bk d_ik)
    const R a = fcik * dfcjk * fcij;
    const R bjj = fcik * dfcij * fcjk, bkk = fcij * dfcik * fcjk;
    const R S2F = S2 * F, S3F = S3 * F;
    const R beta = S1 * a - S2F;
    const R cjk = beta - S3F * inv;
    const R cjj = S2F - S1 * bjj + S3F * cs * irj * irj - beta;
    const R ckk = S2F - S1 * bkk + S3F * cs * irk * irk - beta;
    __hip_atomic_fetch_add(&L.gx[j], (double)(cjj * aj.x + cjk * ak.x), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __hip_atomic_fetch_add(&L.gy[j], (double)(cjj * aj.y + cjk * ak.y), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __hip_atomic_fetch_add(&L.gz[j], (double)(cjj * aj.z + cjk * ak.z), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __hip_atomic_fetch_add(&L.gx[k], (double)(ckk * ak.x + cjk * aj.x), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __hip_atomic_fetch_add(&L.gy[k], (double)(ckk * ak.y + cjk * aj.y), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __hip_atomic_fetch_add(&L.gz[k], (double)(ckk * ak.z + cjk * aj.z), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// The LOGF terms of one function (FP32, lambda = +-1): b^(zeta-1) exp(-eta
// R2) as one exponential of (zeta - 1) log2(b) + ea R2, with log2(1 +- cos)
// once per triplet, no branch on zeta. p = (exp scale, lambda, q1, q2), zm =
// zeta - 1, q3: the weights of S1, S2, S3.
__device__ __forceinline__ void logf_terms(const Vec4<float>::T p, float zm, float q3, float r2, float bp, float bm,
                                           float Lp, float Lm, float &S1, float &S2, float &S3)
{
#pragma clang fp contract(off)
    const bool pos = p.y > 0.0f;
    const float EB = __builtin_amdgcn_exp2f(fmaR(zm, pos ? Lp : Lm, p.x * r2));   // b^(zeta-1) E
    const float C = EB * (pos ? bp : bm);                                        // b^zeta E
    S1 = fmaR(C, p.z, S1);
    S2 = fmaR(C, p.w, S2);
    S3 = fmaR(EB, q3, S3);
}

// Pass B over a chunk of nv triplets in L.pairs, the functions of each
// triplet's block from the LDS tables. LOGF: FP32 with every lambda = +-1
// (logf_terms).
template <typename R, int CT, bool LOGF = false>
__device__ __forceinline__ void g4_grad_list(const DevModel &m, const Lds<R> &L, int nv, int g4s, R rrc)
{
    const int lane = lane_id();
    // each lane takes a contiguous range of the chunk: for a list in block
    // order the lanes of one step then hold different neighbours j and k,
    // whose gradient rows they update
    const int per = (nv + 63) >> 6;
    for (int st = 0; st < per; ++st) {
        const int q = lane * per + st;
        if (q >= nv) break;
        const int pk = L.pairs[q];
        const int j = pk & 0x3FF, k = (pk >> 10) & 0x3FF;
        const I2 bt = L.btab[pk >> 20];
        g4_grad_triplet<R, CT>(m, L, j, k, rrc, [&](const typename Lds<R>::V &aj, const typename Lds<R>::V &ak, R cs,
                                                    R r2, R &S1, R &S2, R &S3) {
            if constexpr (LOGF) {
                float bp, bm;
                one_pm_cos(aj, ak, bp, bm);
                const float Lp = log2_base(bp), Lm = log2_base(bm);
                for (int t = 0; t < bt.y; ++t) {
                    const int u = bt.x + t;
                    logf_terms(L.sp0[u], __int_as_float(L.smeta[u]), L.sp1[u], r2, bp, bm, Lp, Lm, S1, S2, S3);
                }
            } else {
                R E = 0;
                for (int t = 0; t < bt.y; ++t) {
                    const int u = bt.x + t;
                    const typename Lds<R>::V p = L.sp0[u];
                    const int mt = L.smeta[u];
                    if (t == 0 || (mt & (1 << 8))) E = M<R>::expsc(p.x, r2);
                    R b = R(1) + p.y * cs;
                    R bz, bm;
                    const int sq = mt & 0xFF;
                    if (sq != 0xFF) {
                        bm = R(1);
                        for (int kk = 0; kk < sq; ++kk) {
                            bm *= b;
                            b *= b;
                        }
                        bz = b;
                    } else {
                        const int x = g4s + u;
                        bm = angpow<R>(b, cld(m.g4_ipow, x), tabc<R>(m.g4_zeta, m.g4f_zeta, x) - R(1));
                        bz = bm * b;
                    }
                    const R C = E * bz;
                    S1 += C * p.z;
                    S2 += C * p.w;
                    S3 += E * bm * L.sp1[u];
                }
            }
        });
    }
}

// Pass B (FP32, LOGF) over the atom's triplet list in block order, as the
// values pass left it: tb = (entries | 16-bit list << 16, then start | count
// << 11 | block << 23 each). The lanes are dealt out to the blocks in proportion to their
// triplets; a lane keeps its block's function weights in registers.
template <int CT>
__device__ __forceinline__ void g4_grad_blocks(const DevModel &m, const Lds<float> &L, const int *tl, const int *tb,
                                               int V, float rrc)
{
    const int lane = lane_id();
    const int nw = uni(tb[0]) & 0xFFFF;
    const bool f16 = (uni(tb[0]) >> 16) & 1;   // 16-bit entries j | k << 8
    const unsigned short *tl16 = (const unsigned short *)tl;
    const int ent = lane < nw ? tb[1 + lane] : 0;
    const int cx = (ent >> 11) & 0xFFF;
    int T, nlx;
    deal_lanes(V, nw, cx, T, nlx);
    int ws = 0, wc = 0, nl = 1, q = 0, blk = 0;
    {
        int first = 0;
        for (int x = 0; x < nw; ++x) {
            const int n_l = __builtin_amdgcn_readlane(nlx, x), ex = __builtin_amdgcn_readlane(ent, x);
            if (lane >= first && lane < first + n_l) {
                ws = ex & 0x7FF;
                wc = (ex >> 11) & 0xFFF;
                blk = (unsigned)ex >> 23;
                nl = n_l;
                q = lane - first;
            }
            first += n_l;
        }
    }
    const I2 bt = L.btab[blk];
    const int u0 = bt.x, cnt = wc > 0 ? bt.y : 0;
    Vec4<float>::T pp[kRowFns];
    float zm[kRowFns], q3[kRowFns];
#pragma unroll
    for (int t = 0; t < kRowFns; ++t) {
        pp[t] = t < cnt ? L.sp0[u0 + t] : Vec4<float>::T{};
        zm[t] = t < cnt ? __int_as_float(L.smeta[u0 + t]) : 0.0f;
        q3[t] = t < cnt ? L.sp1[u0 + t] : 0.0f;
    }
    auto entry = [&](int x) { return f16 ? (int)tl16[x] : tl[x]; };
    int pk = q < wc ? entry(ws + q) : 0;
    for (int st = 0; st < T; ++st) {
        const int idx = q + st * nl;
        const int cur = pk;
        if (idx + nl < wc) pk = entry(ws + idx + nl);   // the next step's triplet
        if (idx >= wc) continue;
        g4_grad_triplet<float, CT>(m, L, f16 ? cur & 0xFF : cur & 0x3FF, f16 ? cur >> 8 : (cur >> 10) & 0x3FF, rrc,
                                   [&](const Vec4<float>::T &aj, const Vec4<float>::T &ak, float, float r2, float &S1,
                                       float &S2, float &S3) {
                                       float bp, bm;
                                       one_pm_cos(aj, ak, bp, bm);
                                       const float Lp = log2_base(bp), Lm = log2_base(bm);
#pragma unroll
                                       for (int t = 0; t < kRowFns; ++t) {
                                           if (t >= cnt) break;
                                           logf_terms(pp[t], zm[t], q3[t], r2, bp, bm, Lp, Lm, S1, S2, S3);
                                       }
                                       for (int t = kRowFns; t < cnt; ++t)   // beyond the registers'
                                           logf_terms(L.sp0[u0 + t], __int_as_float(L.smeta[u0 + t]), L.sp1[u0 + t],
                                                      r2, bp, bm, Lp, Lm, S1, S2, S3);
                                   });
    }
}

// the tanh^3 cutoff (the reference's) inlined, other types through the switch
template <typename R>
__device__ __forceinline__ void g4_grad_list_ct(const DevModel &m, const Lds<R> &L, int nv, int g4s, R rrc)
{
    if constexpr (sizeof(R) == 4) {
        if (m.g4_lampm) {
            if (m.cutoff_type == 2) g4_grad_list<R, 2, true>(m, L, nv, g4s, rrc);
            else g4_grad_list<R, -1, true>(m, L, nv, g4s, rrc);
            return;
        }
    }
    if (m.cutoff_type == 2) g4_grad_list<R, 2>(m, L, nv, g4s, rrc);
    else g4_grad_list<R, -1>(m, L, nv, g4s, rrc);
}

// ---------------------------------------------------------------- pass A
// Waves per SIMD the FP64 passes at the 96-neighbour staging, one atom per
// wavefront, are compiled for: 3. Without the bound the compiler keeps
// exp_core's 13 constants in VGPRs and lands at 2. The values pass fits in
// the 168 VGPRs of 3 waves with no scratch (at 4, which its LDS allows on
// three SIMDs of four, it spilled 40 bytes per lane, 145 MB of writes per
// 49,152 atoms, for 6% less time); the gradient pass spills 60 bytes per
// lane outside the triplet loops (at 2 waves it spills nothing but takes
// 20% longer).
template <typename R, int NMAX, int GA, bool GRAD>
constexpr int kSfWaves = sizeof(R) == 8 && NMAX <= 96 && GA == 1 ? 3 : 1;

// ALLP (one atom per wavefront): G4 over all pairs of the atom, in pass B's
// order, with the function sums in FP64 (g4_values_list); the compacted
// triplets go to the atom's list for pass B (up to tcap; the count is
// stored in full, and pass B compacts again when it is larger). Else per
// element-pair block (g4_blocks).
template <typename R, int NMAX, int GA, bool ALLP>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(kSfWaves<R, NMAX, GA, false>))) void k_sf_values(DevModel m, Geo g, Work w, R *G)
{
    constexpr int GL = 64 / GA;
    extern __shared__ __align__(16) char smem[];
    int e, t0, tn;
    elem_tile(w, m.nel, blockIdx.x, GA, e, t0, tn);
    if (e < 0) return;
    const int lane = lane_id(), grp = lane / GL, gl = lane % GL;
    const int i = tn < 0 ? t0 : (grp < tn ? w.perm[t0 + grp] : -1);
    const ElemInfo ei = cld_einfo(m.einfo, e);
    const int nsf = ei.nsf;
    const int g4x = ALLP ? m.g4_max : 0;
    const Lds<R> L = make_lds<R>(NMAX, m.nsfmax, g4x, m.nel, false,
                                 (LDS char *)smem + grp * lds_offsets<R>(NMAX, m.nsfmax, g4x, m.nel, false, ALLP).total,
                                 ALLP);
    for (int s = gl; s < nsf; s += GL) {
        if (ALLP) L.accd[s] = 0;
        else L.acc[s] = 0;
    }
    stage<R, GL, NMAX, ALLP && sizeof(R) == 4>(m, g, w, i, wave_max(i >= 0 ? w.count[i] : 0), ei.rc, L, (R *)nullptr);
    __syncthreads();
    const R rce = (R)ei.rc;

    // G2, symmetryFunctions.f90:220-230: one lane per SF of the centre element
    for (int t = gl; t < ei.g2_count; t += GL) {
        const int x = ei.g2_start + t;
        const int f = cld(m.g2_f, x);
        const R ea = g2_ea<R>(m, x), rs = tabc<R>(m.g2_rs, m.g2f_rs, x);
        R sum = 0;
        for (int q = L.seg[f]; q < L.seg[f + 1]; ++q) {
            const typename Lds<R>::V b = L.b[q];
            const R dr = b.x - rs;
            sum += M<R>::expsc(ea, dr * dr) * b.y;
        }
        if (ALLP) L.accd[cld(m.g2_s, x)] += (double)sum;
        else L.acc[cld(m.g2_s, x)] += sum;
    }
    // G4, symmetryFunctions.f90:231-305
    if constexpr (ALLP) {
        values_tables<R>(m, L, e, ei.g4_start, ei.g4_count);
        __syncthreads();
        const int n = uni(L.seg[kMaxEl]), np = n * (n - 1) / 2;
        const float inv_n = 1.0f / (float)max(n, 1);
        const R rrc = R(1) / rce;
        int *tl = (GA == 1 && w.trip && i >= 0) ? w.trip + (size_t)i * w.tcap : nullptr;
        int tbase = 0;
        for (int p0 = 0; p0 < np;) {
            const int nv = uni(compact_circ_fill<R>(L, p0, np, kAllPairsChunk, n, inv_n, rce * rce, m.nel));
            __syncthreads();
            if (tl) {
                const bool t16 = trip16(n);
                for (int q = gl; q < nv && tbase + q < w.tcap; q += 64) {
                    const int pk = L.pairs[q];
                    if (t16) ((unsigned short *)tl)[tbase + q] = (unsigned short)((pk & 0xFF) | ((pk >> 10) & 0xFF) << 8);
                    else tl[tbase + q] = pk;
                }
            }
            tbase += nv;
            if (m.cutoff_type == 2) g4_values_list<R, 2>(m, L, nv, ei.g4_start, rrc);   // tanh^3 inlined
            else g4_values_list<R, -1>(m, L, nv, ei.g4_start, rrc);
            __syncthreads();
        }
        if (tl && gl == 0) w.tcount[i] = tbase;
    } else {
        int *tl = (sizeof(R) == 8 && GA == 1 && w.trip) ? w.trip + (size_t)i * w.tcap : nullptr;
        const int ntrip = g4_blocks<R, GL, false>(m, L, e, rce, tl, w.tcap);
        if (sizeof(R) == 8 && tl && gl == 0) w.tcount[i] = ntrip;
    }
    __syncthreads();
    // centre and scale, hdnnps.f90:471-481
    if (i < 0) return;
    const int so = ei.sf_off;
    for (int s = gl; s < m.nsfmax; s += GL)
        G[(size_t)i * m.nsfmax + s] =
            s < nsf ? (R)(((ALLP ? L.accd[s] : (double)L.acc[s]) - m.sf_mean[so + s]) / m.sf_den[so + s]) : R(0);
}

// ---------------------------------------------------------------- pass B
template <typename R, int NMAX, int GA>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(kSfWaves<R, NMAX, GA, true>))) void k_sf_grad(DevModel m, Geo g, Work w, const R *dEdG, R *gent)
{
    constexpr int GL = 64 / GA;
    extern __shared__ __align__(16) char smem[];
    int e, t0, tn;
    elem_tile(w, m.nel, blockIdx.x, GA, e, t0, tn);
    if (e < 0) return;
    const int lane = lane_id(), grp = lane / GL, gl = lane % GL;
    const int i = tn < 0 ? t0 : (grp < tn ? w.perm[t0 + grp] : -1);
    const ElemInfo ei = cld_einfo(m.einfo, e);
    const int nsf = ei.nsf;
    const int so = ei.sf_off;
    const Lds<R> L = make_lds<R>(NMAX, m.nsfmax, m.g4_max, m.nel, true,
                                 (LDS char *)smem + grp * lds_offsets<R>(NMAX, m.nsfmax, m.g4_max, m.nel, true).total);
    // dE/dG of the unscaled G: the scaled G is (G - mean)/den
    for (int s = gl; s < nsf; s += GL)
        L.acc[s] = i >= 0 ? (R)((double)dEdG[(size_t)i * m.nsfmax + s] / m.sf_den[so + s]) : R(0);
    stage<R, GL, NMAX, true>(m, g, w, i, wave_max(i >= 0 ? w.count[i] : 0), ei.rc, L, gent);
    __syncthreads();
    const int n = gu<GL>(L.seg[kMaxEl]);
    for (int q = gl; q < n; q += GL) L.gx[q] = L.gy[q] = L.gz[q] = 0;
    __syncthreads();
    const R rce = (R)ei.rc;

    // G2: g_j = d_ij * (-dfc * sum w E + fc/r * sum w 2 eta (r - Rs) E), one lane per neighbour
    for (int q = gl; q < n; q += GL) {
        int tbx, tby;
        cld2(m.g2_tab, e * kMaxEl + L.el[q], tbx, tby);
        const typename Lds<R>::V a = L.a[q], b = L.b[q];
        R A = 0, B = 0;
        for (int t = 0; t < tby; ++t) {
            const int x = tbx + t;
            const R dr = b.x - tabc<R>(m.g2_rs, m.g2f_rs, x);
            const R ws = L.acc[cld(m.g2_s, x)] * M<R>::expsc(g2_ea<R>(m, x), dr * dr);
            A += ws;
            B += ws * R(2) * tabc<R>(m.g2_eta, m.g2f_eta, x) * dr;
        }
        const R c = -b.z * A + b.y * B * a.w;
        L.gx[q] += (double)(c * a.x);
        L.gy[q] += (double)(c * a.y);
        L.gz[q] += (double)(c * a.z);
    }
    __syncthreads();
    // G4
    if (GA == 1) {
        grad_tables<R>(m, L, e, ei.g4_start, ei.g4_count, sizeof(R) == 4 && m.g4_lampm);
        __syncthreads();
        const R rce2 = rce * rce, rrc = R(1) / rce;
        const int tc = w.trip ? uni(w.tcount[i]) : -1;
        bool done = false;
        if constexpr (sizeof(R) == 4) {
            if (tc >= 0 && tc <= w.tcap && w.tblk && m.g4_lampm) {   // the block-ordered list of the FP32 pass
                const int *tl = w.trip + (size_t)i * w.tcap, *tb = w.tblk + (size_t)i * kTripBlk;
                if (m.cutoff_type == 2) g4_grad_blocks<2>(m, L, tl, tb, tc, rrc);
                else g4_grad_blocks<-1>(m, L, tl, tb, tc, rrc);
                done = true;
            }
        }
        if (done) {
        } else if (tc >= 0 && tc <= w.tcap) {
            // the triplets pass A compacted
            const int *tl = w.trip + (size_t)i * w.tcap;
            const bool t16 = trip16(n);
            for (int cb = 0; cb < tc; cb += kPairChunk<R>) {
                const int nv = min(kPairChunk<R>, tc - cb);
                for (int q = gl; q < nv; q += 64) {
                    if (t16) {   // the block from the two neighbours' elements
                        const int e16 = ((const unsigned short *)tl)[cb + q];
                        const int j = e16 & 0xFF, k = e16 >> 8, ej = L.el[j], ek = L.el[k];
                        L.pairs[q] = j | k << 10 | (min(ej, ek) * m.nel + max(ej, ek)) << 20;
                    } else {
                        L.pairs[q] = tl[cb + q];
                    }
                }
                __syncthreads();
                g4_grad_list_ct<R>(m, L, nv, ei.g4_start, rrc);
                __syncthreads();
            }
        } else {
            const int np = n * (n - 1) / 2;
            const float inv_n = 1.0f / (float)max(n, 1);
            for (int p0 = 0; p0 < np;) {
                const int nv = uni(compact_circ_fill<R>(L, p0, np, kPairChunk<R>, n, inv_n, rce2, m.nel));
                __syncthreads();
                g4_grad_list_ct<R>(m, L, nv, ei.g4_start, rrc);
                __syncthreads();
            }
        }
    } else {
        g4_blocks<R, GL, true>(m, L, e, rce);
    }
    // per-entry gradient rows and the atom's virial -sum_n g_n (x) d_in
    double v[9];
#pragma unroll
    for (int t = 0; t < 9; ++t) v[t] = 0;
    for (int q = gl; q < n; q += GL) {
        const double gx = L.gx[q], gy = L.gy[q], gz = L.gz[q];
        const size_t en = (size_t)i * w.cap + L.ent[q];
        gent[3 * en + 0] = (R)gx;
        gent[3 * en + 1] = (R)gy;
        gent[3 * en + 2] = (R)gz;
        const typename Lds<R>::V a = L.a[q];
        const double dxq = a.x, dyq = a.y, dzq = a.z;
        v[0] -= gx * dxq; v[1] -= gx * dyq; v[2] -= gx * dzq;
        v[3] -= gy * dxq; v[4] -= gy * dyq; v[5] -= gy * dzq;
        v[6] -= gz * dxq; v[7] -= gz * dyq; v[8] -= gz * dzq;
    }
#pragma unroll
    for (int t = 0; t < 9; ++t) {
        const double s = group_sum<GL>(v[t]);
        if (gl == 0 && i >= 0) w.vir[9 * (size_t)i + t] = s;
    }
}

// ---------------------------------------------------------------- FP32 block-sweep values pass
// Pass A in FP32 with one atom per wavefront when every G4 lambda is +-1 (so
// 1 + lambda cos >= 0) and every zeta >= 1, in two phases over the atom's
// element-pair blocks (staged neighbours are grouped by element, so a block
// is a segment pair).
// Phase 1 sweeps each block's candidate pairs as rows (row a of the smaller
// segment against the columns of the other, for one element the circulant
// rows of its segment; the row cut into chunks, a lane per (row, chunk)):
// the lane keeps its row's vector in registers, and at each step the lanes
// read distinct columns (no LDS bank conflicts). The pairs with r_jk < rc
// (about half of the candidates) go to the atom's triplet list, compacted by
// ballot, in block order; the block ranges go to a small table.
// Phase 2 deals the lanes out to the blocks in proportion to their triplets
// (deal_lanes): a lane keeps its block's functions (up to kRowFns) and its
// term sums in registers, sums kSumRun terms in FP32 and then in FP64, and
// adds them to the FP64 function sums (LDS atomics) once at the end -- not
// one LDS update per term. The angular and radial factors of a function are
// one exponential:
//   2^(1-zeta) (1 + lambda cos)^zeta exp(-eta R2) = exp2(zeta L + (1 - zeta) + ea R2)
// with L = log2(1 + lambda cos) (two logarithms per triplet, one per sign of
// lambda; 1 +- cos from the unit vectors, one_pm_cos; L >= -126, so b = 0
// gives 0 for zeta > 1), R2 = r_ij^2 + r_ik^2 + r_jk^2 and ea = -eta
// log2(e): no branch on zeta, and any real zeta >= 1. Functions of a block
// beyond kRowFns read their parameters from LDS per triplet.
// The list holds tcap entries per atom; an atom with more triplets is
// evaluated in windows of the list (phase 1 until the window is full, then
// phase 2), and pass B then compacts its triplets itself. Pass B reads the
// list and the table of a complete list (g4_grad_blocks). The wavefront of
// atom b runs on XCD b % 8: xcd_block gives each XCD a contiguous range of
// atoms, whose neighbours overlap in its L2.

struct DirectOffsets {
    size_t a, b, el, ent, seg, accd, tab0, blk, win, total;
};

__host__ __device__ __forceinline__ DirectOffsets direct_offsets(int nmax, int nsfmax, int g4max, int nblk)
{
    DirectOffsets o{};
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t at = off;
        off += (bytes + 15) / 16 * 16;
        return at;
    };
    o.a = take(16 * (size_t)nmax);
    o.b = take(16 * (size_t)nmax);
    o.el = take((size_t)nmax);
    o.ent = take(2 * (size_t)nmax);
    o.seg = take(4 * (kMaxEl + 1));
    o.accd = take(8 * (size_t)nsfmax);   // G sums (FP64)
    o.tab0 = take(16 * (size_t)g4max);
    o.blk = take(16 * (size_t)nblk);     // element-pair blocks with G4 functions
    o.win = take(8 * (size_t)nblk);      // the blocks of a window of the triplet list
    o.total = off;
    return o;
}

// the stage() view of the block-sweep pass's LDS
__device__ __forceinline__ Lds<float> direct_lds(const DirectOffsets &o, LDS char *base)
{
    Lds<float> L{};
    L.a = (LDS Vec4<float>::T *)(base + o.a);
    L.b = (LDS Vec4<float>::T *)(base + o.b);
    L.el = (LDS unsigned char *)(base + o.el);
    L.ent = (LDS unsigned short *)(base + o.ent);
    L.seg = (LDS int *)(base + o.seg);
    L.accd = (LDS double *)(base + o.accd);
    return L;
}

// log2(b) of b = 1 + lambda cos, b clamped to [0, ...) against rounding and
// the logarithm to -126 (so that 0 * L = 0)
__device__ __forceinline__ float log2_base(float b)
{
    return __builtin_fmaxf(__builtin_amdgcn_logf(__builtin_fmaxf(b, 0.0f)), -126.0f);
}

template <int NMAX, int CT>
__global__ __launch_bounds__(64) void k_sf_values_direct(DevModel m, Geo g, Work w, float *G)
{
#pragma clang fp contract(off)
    extern __shared__ __align__(16) char smem[];
    const int i = xcd_block(blockIdx.x, w.nat);
    const int lane = lane_id();
    const int e = uni(w.eidx[i]);
    const ElemInfo ei = cld_einfo(m.einfo, e);
    const int nsf = ei.nsf, g4s = ei.g4_start, g4n = ei.g4_count;
    const int ct = CT >= 0 ? CT : m.cutoff_type;
    const DirectOffsets o = direct_offsets(NMAX, m.nsfmax, m.g4_max, m.g4_nblk);
    const Lds<float> L = direct_lds(o, (LDS char *)smem);
    LDS Vec4<float>::T *tab = (LDS Vec4<float>::T *)((LDS char *)smem + o.tab0);
    for (int s = lane; s < nsf; s += 64) L.accd[s] = 0;
    // (ea, zeta, lambda, G slot) of each local G4 function
    for (int u = lane; u < g4n; u += 64) {
        const int x = g4s + u;
        tab[u] = Vec4<float>::T{cld(m.g4f_ea, x), cld(m.g4f_zeta, x), cld(m.g4f_lam, x),
                                __int_as_float(cld(m.g4_s, x))};
    }
    stage<float, 64, NMAX, true>(m, g, w, i, wave_max(w.count[i]), ei.rc, L, (float *)nullptr);
    __syncthreads();
    const float rce = (float)ei.rc;

    // G2, symmetryFunctions.f90:220-230: one lane per SF of the centre element
    for (int t = lane; t < ei.g2_count; t += 64) {
        const int x = ei.g2_start + t;
        const int f = cld(m.g2_f, x);
        const float ea = g2_ea<float>(m, x), rs = cld(m.g2f_rs, x);
        float sum = 0;
        for (int q = L.seg[f]; q < L.seg[f + 1]; ++q) {
            const Vec4<float>::T b = L.b[q];
            const float dr = b.x - rs;
            sum = fmaR(M<float>::expsc(ea, dr * dr), b.y, sum);
        }
        L.accd[cld(m.g2_s, x)] += (double)sum;
    }
    __syncthreads();
    // G4, symmetryFunctions.f90:231-305
    LDS I4 *bt = (LDS I4 *)((LDS char *)smem + o.blk);
    LDS I2 *wt = (LDS I2 *)((LDS char *)smem + o.win);
    const int nel = m.nel;
    // The element-pair blocks (f <= h) with G4 functions, in (f, h) order,
    // each swept as rows: row a of segment (sa, na) against the columns of
    // segment (sb, nb), the row cut into C chunks of len columns, na C items
    // in passes of 64 lanes. Two elements: (sa, na) is the smaller segment
    // and row a takes the columns (a + o) mod nb, o < nb (diagonals: the
    // lanes of a chunk read different columns). One element (n = na = nb):
    // row a takes (a + 1 + o) mod n, o < n/2, and for even n the last (o =
    // n/2 - 1) only for a < n/2. C fills the 64 lanes: min(R, 64 / na).
    // Entry: (sa | na << 16, sb | nb << 16, first function | count << 16,
    // f nel + h | one element << 8 | C << 9 | len << 16)
    int nbk = 0;
    for (int t0 = 0; t0 < nel * nel; t0 += 64) {
        const int t = t0 + lane;
        const int f = t / nel, h = t - f * nel;
        int bx = 0, by = 0;
        if (t < nel * nel && f <= h) cld2(m.g4_tab, (e * kMaxEl + f) * kMaxEl + h, bx, by);
        const unsigned long long mk = __ballot(by > 0);
        if (by > 0) {
            const int sf = L.seg[f], Lf = L.seg[f + 1] - sf, sh = L.seg[h], Lh = L.seg[h + 1] - sh;
            const bool same = f == h;
            const int sa = same || Lf <= Lh ? sf : sh, na = min(Lf, Lh);
            const int sb = same || Lf <= Lh ? sh : sf, nb = max(Lf, Lh);
            const int R = same ? na / 2 : nb;
            const int C = max(1, na >= 64 ? 1 : min(R, 64 / max(na, 1)));
            const int len = (R + C - 1) / C;
            bt[nbk + mbcnt(mk)] = I4{sa | na << 16, sb | nb << 16, (bx - g4s) | by << 16,
                                     t | (int)same << 8 | C << 9 | len << 16};
        }
        nbk += __popcll(mk);
    }
    __syncthreads();
    const float rrc = 1.0f / rce, rc2 = rce * rce;
    int *tl = w.trip + (size_t)i * w.tcap;
    unsigned short *tl16 = (unsigned short *)tl;   // the flat schedule's entries
    const int cap = w.tcap;
    int b = 0, P = 0, sp = 0, tot = 0, nw = 0;   // the sweep's block, pass and step
    bool over = false;
    // Flat schedule, when the list has a place for every candidate pair and
    // pass B reads it through the block table: block b gets the list region
    // [base_b, base_b + np_b), and the items (row, chunk of at most
    // kFlatLen columns) of all blocks, in block order, run in passes of 64
    // lanes (passes mix blocks; no block ends a pass early). The triplets of
    // a step go to their blocks' regions, ranked within the block's lanes.
    // The entries are 16-bit, j | k << 8 (the block is the region's): the
    // list an atom writes and reads back is half the size, and more of it is
    // still in L2 when phase 2 reads it.
    // wt[b] = (base | first item << 16, triplets so far | b << 16); bt[b].w
    // takes the flat chunks and length.
    bool flat = false;
    int items_all = 0;
    if (w.tblk) {
        int items_l = 0, np_l = 0, Cf = 1, lenf = 0;
        if (lane < nbk) {
            const I4 d = bt[lane];
            const int na = d.x >> 16, nb = d.y >> 16;
            const bool same = (d.w >> 8) & 1;
            const int R = same ? na / 2 : nb;
            Cf = max(1, (R + kFlatLen - 1) / kFlatLen);
            lenf = (R + Cf - 1) / Cf;
            items_l = na > 0 && R > 0 ? na * Cf : 0;
            np_l = same ? na * (na - 1) / 2 : na * nb;
        }
        const int items_in = wave_scan_incl(items_l), np_in = wave_scan_incl(np_l);
        items_all = __builtin_amdgcn_readlane(items_in, 63);
        flat = __builtin_amdgcn_readlane(np_in, 63) <= cap && uni(L.seg[kMaxEl]) <= 256;   // 8-bit j and k
        if (flat && lane < nbk) {
            bt[lane].w = (bt[lane].w & 0x1FF) | Cf << 9 | lenf << 16;
            wt[lane] = I2{(np_in - np_l) | (items_in - items_l) << 16, lane << 16};
        }
        __syncthreads();
    }
    while (b < nbk) {
        // phase 1: the triplets with r_jk < rc of the blocks, in block order,
        // into the list (a window of it: the list holds cap entries)
        bool full = false;
        tot = 0;
        nw = 0;
        if (flat) {
            for (int P0 = 0; P0 < items_all; P0 += 64) {
                const int it = P0 + lane;
                int bb = 0;   // the item's block: the last with first item <= it
#pragma unroll
                for (int st = 32; st > 0; st >>= 1)
                    if (bb + st < nbk && ((int)wt[bb + st].x >> 16) <= it) bb += st;
                const I4 d = bt[bb];
                const I2 wx = wt[bb];
                const int sa = d.x & 0xFFFF, na = d.x >> 16, sb = d.y & 0xFFFF, nb = d.y >> 16;
                const bool same = (d.w >> 8) & 1;
                const int Cf = (d.w >> 9) & 0x7F, lenf = d.w >> 16;
                const int first = (int)wx.x >> 16, base = (int)wx.x & 0xFFFF;
                const int R = same ? na / 2 : nb, nitems = na * Cf;
                const bool live = it < items_all;
                const int loc = it - first;
                const int c = row_of(loc, __builtin_amdgcn_rcpf((float)max(na, 1))), a = loc - __mul24(c, na);
                const int o0 = c * lenf;
                int oend = live ? min(R, o0 + lenf) : 0;
                if (same && (na & 1) == 0 && a >= na / 2) oend = min(oend, R - 1);
                const int j = sa + a;
                const Vec4<float>::T aj = L.a[j];
                int kk = a + o0 + (int)same;
                if (kk >= nb) kk -= nb;
                if (kk >= nb) kk -= nb;
                if (!live || kk >= nb) kk = 0;
                // the lanes of this block in the pass
                const int lo = max(first - P0, 0), hi = min(first + nitems - P0, 64);
                const unsigned long long bmask =
                    live ? (hi >= 64 ? ~0ull : (1ull << hi) - 1) & ~((1ull << lo) - 1) : 0ull;
                int cnt = (int)wx.y & 0xFFFF;
                LDS const Vec4<float>::T *colp = L.a + sb;
#pragma unroll
                for (int s = 0; s < kFlatLen; ++s) {
                    const Vec4<float>::T ak = colp[kk];
                    const float ex = ak.x - aj.x, ey = ak.y - aj.y, ez = ak.z - aj.z;
                    const bool valid = (o0 + s < oend) & (fmaR(ez, ez, fmaR(ey, ey, ex * ex)) < rc2);
                    const unsigned long long mk = __ballot(valid) & bmask;
                    if (valid) tl16[base + cnt + mbcnt(mk)] = (unsigned short)(j | (sb + kk) << 8);
                    cnt += __popcll(mk);
                    kk = kk + 1 == nb ? 0 : kk + 1;
                }
                __syncthreads();
                if (live && lane == hi - 1) wt[bb].y = cnt | bb << 16;
                __syncthreads();
            }
            b = nbk;
            nw = nbk;
            tot = __builtin_amdgcn_readlane(wave_scan_incl(lane < nbk ? ((int)wt[lane].y & 0xFFFF) : 0), 63);

        } else
        for (; b < nbk; ++b, P = 0, sp = 0) {
            const I4 d = bt[b];
            const int dx = uni(d.x), dy = uni(d.y), dw = uni(d.w);
            const int sa = dx & 0xFFFF, na = dx >> 16, sb = dy & 0xFFFF, nb = dy >> 16;
            const bool same = (dw >> 8) & 1;
            const int C = (dw >> 9) & 0x7F, len = dw >> 16, code = (dw & 0xFF) << 20;
            const int R = same ? na / 2 : nb, items = na * C;
            const float inv_na = __builtin_amdgcn_rcpf((float)max(na, 1));
            const int bstart = tot;
            for (; P * 64 < items; ++P, sp = 0) {
                const int it = P * 64 + lane;
                const int c = row_of(it, inv_na), a = it - __mul24(c, na);
                const int o0 = c * len;
                // columns o0 + sp .. oend - 1 of row a (kk, the column, stays
                // in [0, nb) also on lanes without columns: their reads stay
                // in the segment)
                int oend = it < items ? min(R, o0 + len) : 0;
                if (same && (na & 1) == 0 && a >= na / 2) oend = min(oend, R - 1);
                const int j = sa + a;
                const Vec4<float>::T aj = L.a[j];
                int kk = a + o0 + sp + (int)same;
                if (kk >= nb) kk -= nb;
                if (kk >= nb) kk -= nb;
                if (kk >= nb) kk = 0;
                const int jc = j | code;
                LDS const Vec4<float>::T *colp = L.a + sb;
                // the pass writes at most (its lanes) x (its steps left)
                // entries: when they fit, its steps skip the capacity test
                if (tot + min(64, items - P * 64) * (len - sp) <= cap) {
                    for (int o = o0 + sp; sp < len; ++sp, ++o) {
                        const Vec4<float>::T ak = colp[kk];
                        const float ex = ak.x - aj.x, ey = ak.y - aj.y, ez = ak.z - aj.z;
                        const bool valid = (o < oend) & (fmaR(ez, ez, fmaR(ey, ey, ex * ex)) < rc2);
                        const unsigned long long mk = __ballot(valid);
                        if (valid) tl[tot + mbcnt(mk)] = jc | ((sb + kk) << 10);
                        tot += __popcll(mk);
                        kk = kk + 1 == nb ? 0 : kk + 1;
                    }
                } else {
                    for (int o = o0 + sp; sp < len; ++sp, ++o) {
                        const Vec4<float>::T ak = colp[kk];
                        const float ex = ak.x - aj.x, ey = ak.y - aj.y, ez = ak.z - aj.z;
                        const bool valid = (o < oend) & (fmaR(ez, ez, fmaR(ey, ey, ex * ex)) < rc2);
                        const unsigned long long mk = __ballot(valid);
                        const int cv = __popcll(mk);
                        if (tot > 0 && tot + cv > cap) {   // the window is full: evaluate it, then go on
                            full = true;
                            break;
                        }
                        const int pos = tot + mbcnt(mk);
                        if (valid && pos < cap) tl[pos] = jc | ((sb + kk) << 10);
                        tot += cv;
                        kk = kk + 1 == nb ? 0 : kk + 1;
                    }
                    if (full) break;
                }
            }
            const int te = min(tot, cap);
            if (te > bstart) {
                if (lane == 0) wt[nw] = I2{bstart, (te - bstart) | (b << 16)};
                ++nw;
            }
            if (full) break;
        }
        __syncthreads();
        // phase 2: the lanes are dealt out to the window's blocks in
        // proportion to their triplets; a lane keeps its block's functions
        // and its term sums in registers
        const int V = min(tot, cap);
        int T;
        {   // lane x < nw: the lanes of entry x
            const int cx = lane < nw ? ((int)wt[lane].y & 0xFFFF) : 0;
            int nlx;
            deal_lanes(V, nw, cx, T, nlx);
            if (lane < nw) wt[lane].x = ((int)wt[lane].x & 0xFFFF) | nlx << 16;
        }
        __syncthreads();
        int ws = 0, wc = 0, nl = 1, q = 0, u0 = 0, cnt = 0;
        {
            int first = 0;
            for (int x = 0; x < nw; ++x) {
                const I2 wx = wt[x];
                const int wx0 = uni((int)wx.x), wx1 = uni((int)wx.y);
                const int n_l = wx0 >> 16;
                if (lane >= first && lane < first + n_l) {
                    ws = wx0 & 0xFFFF;
                    wc = wx1 & 0xFFFF;
                    nl = n_l;
                    q = lane - first;
                    const int bz = bt[wx1 >> 16].z;
                    u0 = bz & 0xFFFF;
                    cnt = bz >> 16;
                }
                first += n_l;
            }
        }
        // the term sums: FP32 over (at most) kSumRun steps, then FP64
        float ea[kRowFns], ze[kRowFns], cz[kRowFns], lam[kRowFns], acc[kRowFns];
        double accd[kRowFns];
        int slot[kRowFns];
#pragma unroll
        for (int t = 0; t < kRowFns; ++t) {
            const Vec4<float>::T P = t < cnt ? tab[u0 + t] : Vec4<float>::T{};
            ea[t] = P.x;
            ze[t] = P.y;
            cz[t] = 1.0f - P.y;
            lam[t] = P.z;
            slot[t] = __float_as_int(P.w);
            acc[t] = 0;
            accd[t] = 0;
        }
        auto entry = [&](int x) { return flat ? (int)tl16[x] : tl[x]; };
        int pk = q < wc ? entry(ws + q) : 0;
        for (int st = 0; st < T; ++st) {
            if (st % kSumRun == 0 && st > 0) {
#pragma unroll
                for (int t = 0; t < kRowFns; ++t) {
                    accd[t] += (double)acc[t];
                    acc[t] = 0;
                }
            }
            const int idx = q + st * nl;
            const int cur = pk;
            if (idx + nl < wc) pk = entry(ws + idx + nl);   // the next step's triplet
            if (idx >= wc) continue;
            const int j = flat ? cur & 0xFF : cur & 0x3FF, k = flat ? cur >> 8 : (cur >> 10) & 0x3FF;
            const Vec4<float>::T aj = L.a[j], ak = L.a[k], bj = L.b[j], bk = L.b[k];
            const float ex = ak.x - aj.x, ey = ak.y - aj.y, ez = ak.z - aj.z;
            const float rjk2 = fmaR(ez, ez, fmaR(ey, ey, ex * ex));
            const float fcjk = cutoff_in<float>(ct, M<float>::sqrt(rjk2), rrc);
            float bp, bm;
            one_pm_cos(aj, ak, bp, bm);
            const float F = (bj.y * bk.y) * fcjk;
            const float r2 = (bj.w + bk.w) + rjk2;
            const float Lp = log2_base(bp), Lm = log2_base(bm);
#pragma unroll
            for (int t = 0; t < kRowFns; ++t) {
                if (t >= cnt) break;
                const float Ls = lam[t] > 0.0f ? Lp : Lm;
                acc[t] = fmaR(__builtin_amdgcn_exp2f(fmaR(ze[t], Ls, fmaR(ea[t], r2, cz[t]))), F, acc[t]);
            }
            for (int t = kRowFns; t < cnt; ++t) {   // functions beyond the registers' (not in common models)
                const Vec4<float>::T P = tab[u0 + t];
                const float Ls = P.z > 0.0f ? Lp : Lm;
                const float Tv = __builtin_amdgcn_exp2f(fmaR(P.y, Ls, fmaR(P.x, r2, 1.0f - P.y))) * F;
                __hip_atomic_fetch_add(&L.accd[__float_as_int(P.w)], (double)Tv, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_WORKGROUP);
            }
        }
#pragma unroll
        for (int t = 0; t < kRowFns; ++t)
            if (t < cnt) __hip_atomic_fetch_add(&L.accd[slot[t]], accd[t] + (double)acc[t], __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_WORKGROUP);
        over = over || full;
        if (!over && w.tblk && lane < nw) {   // the blocks of the whole list, for pass B
            const I2 wx = wt[lane];
            w.tblk[(size_t)i * kTripBlk + 1 + lane] =
                ((int)wx.x & 0x7FF) | ((int)wx.y & 0xFFFF) << 11 | (bt[(int)wx.y >> 16].w & 0xFF) << 23;
        }
        __syncthreads();
    }
    if (lane == 0) {
        w.tcount[i] = over ? cap + 1 : tot;   // cap + 1: pass B compacts its own
        if (w.tblk) w.tblk[(size_t)i * kTripBlk] = over ? 0 : nw | (int)flat << 16;
    }
    __syncthreads();
    // centre and scale, hdnnps.f90:471-481
    const int so = ei.sf_off;
    for (int s = lane; s < m.nsfmax; s += 64)
        G[(size_t)i * m.nsfmax + s] = s < nsf ? (float)((L.accd[s] - m.sf_mean[so + s]) / m.sf_den[so + s]) : 0.0f;
}

bool sf_allpairs_enabled();
int sf_bucket(int nmax);

// the direct FP32 passes serve this model, tile and staging size (both fit
// 64 KiB of LDS)
static bool sf_direct(const DevModel &m, int ga, int nmax)
{
    static const bool on = [] {   // HDNNP_SF_DIRECT=0 in the environment selects the compacting passes
        const char *v = std::getenv("HDNNP_SF_DIRECT");
        return !(v && std::strcmp(v, "0") == 0);
    }();
    return on && ga == 1 && m.g4_lampm && m.g4_nblk <= 48 && sf_allpairs_enabled() &&
           direct_offsets(nmax, m.nsfmax, m.g4_max, m.g4_nblk).total <= 65536;
}

// wavefronts of a launch: tiles of GA atoms of one element
static int sf_blocks(const DevModel &m, const Work &w, int ga) { return ga == 1 ? w.nat : (w.nat + ga - 1) / ga + m.nel; }

// pass A over all pairs with one atom per wavefront, unless
// HDNNP_SF_ALLPAIRS=0 in the environment selects the per-block pass
bool sf_allpairs_enabled()
{
    static const bool on = [] {
        const char *v = std::getenv("HDNNP_SF_ALLPAIRS");
        return !(v && std::strcmp(v, "0") == 0);
    }();
    return on;
}

template <typename R>
void launch_sf_values(const DevModel &m, const Geo &g, Work &w, int nmax, int ga, R *G, hipStream_t s)
{
    const Work wk = w;
    w.tblk = nullptr;   // only the FP32 block-sweep pass writes the block tables
    if constexpr (sizeof(R) == 4) {
        const int nb = sf_bucket(nmax);
        if (sf_direct(m, ga, nb) && w.trip && w.tcount) {   // the pass writes its triplets to the list
            w.tblk = wk.tblk;
            const size_t lds = direct_offsets(nb, m.nsfmax, m.g4_max, m.g4_nblk).total;
#define HDNNP_SFD(N)                                                                                     \
    do {                                                                                                 \
        if (m.cutoff_type == 2)                                                                          \
            hipLaunchKernelGGL((k_sf_values_direct<N, 2>), dim3(w.nat), dim3(64), lds, s, m, g, w, G);   \
        else                                                                                             \
            hipLaunchKernelGGL((k_sf_values_direct<N, -1>), dim3(w.nat), dim3(64), lds, s, m, g, w, G);  \
    } while (0)
            if (w.nat == 0) return;
            if (nb == 64) HDNNP_SFD(64);
            else if (nb == 96) HDNNP_SFD(96);
            else if (nb == 128) HDNNP_SFD(128);
            else if (nb == 256) HDNNP_SFD(256);
            else if (nb == 512) HDNNP_SFD(512);
            else HDNNP_SFD(1024);
#undef HDNNP_SFD
            return;
        }
    }
    const bool allp = ga == 1 && sf_allpairs_enabled();
#define HDNNP_SFV(N, A)                                                                            \
    do {                                                                                           \
        if (allp && A == 1)                                                                        \
            hipLaunchKernelGGL((k_sf_values<R, N, A, A == 1>), dim3(sf_blocks(m, w, A)),            \
                               dim3(64), lds_bytes<R>(N, A, m.nsfmax, m.g4_max, m.nel, false, true), s, m, g, w, G); \
        else                                                                                       \
            hipLaunchKernelGGL((k_sf_values<R, N, A, false>), dim3(sf_blocks(m, w, A)), dim3(64),  \
                               lds_bytes<R>(N, A, m.nsfmax, 0, m.nel, false), s, m, g, w, G);       \
    } while (0)
    if (nmax <= 64) {
        if (ga == 4) HDNNP_SFV(64, 4);
        else if (ga == 2) HDNNP_SFV(64, 2);
        else HDNNP_SFV(64, 1);
    } else if (nmax <= 96 && ga == 1) {
        HDNNP_SFV(96, 1);
    } else if (nmax <= 128) {
        if (ga == 4) HDNNP_SFV(128, 4);
        else if (ga == 2) HDNNP_SFV(128, 2);
        else HDNNP_SFV(128, 1);
    } else if (nmax <= 256) HDNNP_SFV(256, 1);
    else if (nmax <= 512) HDNNP_SFV(512, 1);
    else HDNNP_SFV(1024, 1);
#undef HDNNP_SFV
}

template <typename R>
void launch_sf_grad(const DevModel &m, const Geo &g, Work &w, int nmax, int ga, const R *dEdG, R *gent,
                    hipStream_t s)
{
#define HDNNP_SFG(N, A)                                                                            \
    hipLaunchKernelGGL((k_sf_grad<R, N, A>), dim3(sf_blocks(m, w, A)), dim3(64),                    \
                       lds_bytes<R>(N, A, m.nsfmax, m.g4_max, m.nel, true), s, m, g, w, dEdG, gent)
    if (nmax <= 64) {
        if (ga == 4) HDNNP_SFG(64, 4);
        else if (ga == 2) HDNNP_SFG(64, 2);
        else HDNNP_SFG(64, 1);
    } else if (nmax <= 96 && ga == 1) {
        HDNNP_SFG(96, 1);
    } else if (nmax <= 128) {
        if (ga == 4) HDNNP_SFG(128, 4);
        else if (ga == 2) HDNNP_SFG(128, 2);
        else HDNNP_SFG(128, 1);
    } else if (nmax <= 256) HDNNP_SFG(256, 1);
    else if (nmax <= 512) HDNNP_SFG(512, 1);
    else HDNNP_SFG(1024, 1);
#undef HDNNP_SFG
}

template void launch_sf_values<double>(const DevModel &, const Geo &, Work &, int, int, double *, hipStream_t);
template void launch_sf_values<float>(const DevModel &, const Geo &, Work &, int, int, float *, hipStream_t);
template void launch_sf_grad<double>(const DevModel &, const Geo &, Work &, int, int, const double *, double *,
                                     hipStream_t);
template void launch_sf_grad<float>(const DevModel &, const Geo &, Work &, int, int, const float *, float *,
                                    hipStream_t);

// staging capacity of the kernel instance that serves nmax neighbours
// the staging size of the one-atom-per-wavefront kernels for nmax entries
int sf_bucket(int nmax)
{
    if (nmax <= 64) return 64;
    if (nmax <= 96) return 96;
    int b = 128;
    while (b < nmax) b *= 2;
    return b;
}

// base^y for the angular terms with base >= 0: 0^0 = 1 and 0^y = 0 otherwise
// (the FP32 pow, exp2(y log2 x), gives NaN for 0^0; opposite images of an
// atom in a cell smaller than the cutoff make cos theta = -1 exactly)
template <typename R>
__device__ __forceinline__ R powz(R base, R y)
{
    return base > R(0) ? M<R>::pow(base, y) : (y == R(0) ? R(1) : R(0));
}

// ---------------------------------------------------------------- per-function path
// symmetryFunction2 / symmetryFunction3 (symmetryFunctions.f90:41-152), which
// hdnnpCalcSymfunctions runs with calcSfsTogether = .false.
// (hdnnps.f90:436-467): one function at a time, with dG/dR materialised as
// the reference stores it (per atom, per function, one row per list slot).
// One wavefront per atom. The atom's list entries are staged in LDS as
// d = R_i - R_n, r, fc and fc'/r. Lane q owns entry q (and q + 64, ...): its
// gradient row is the sum over the partners k of the terms of the pairs
// (q, k) it belongs to, so no two lanes write one row and every sum runs in
// a fixed order; the value is half the sum of the lanes' pair sums (each pair
// is seen from both ends); the self row is minus the sum of the other rows
// (translation invariance). Values and rows are centred / scaled as the
// fused pass's.
template <typename R>
__global__ __launch_bounds__(64) void k_sf_single(DevModel m, Geo g, Work w, const SfParam *par, const int *par_off,
                                                  const long long *dsf_off, int ldG, double *Gout, double *dsf)
{
    extern __shared__ __align__(16) char smem[];
    const int i = blockIdx.x, lane = threadIdx.x;
    const int n = w.count[i], e = w.eidx[i];
    R *dx = (R *)smem, *dy = dx + w.lim, *dz = dy + w.lim, *rr = dz + w.lim, *fcs = rr + w.lim, *dfs = fcs + w.lim;
    unsigned char *el = (unsigned char *)(dfs + w.lim);
    const double rce = m.rc[e];
    const R rc = (R)rce, rrc = R(1) / rc;
    const double4 pi = w.wpos[i];
    const Lat9 lt = atom_lattice(g, w, i);
    for (int q = lane; q < n; q += 64) {
        double d0, d1, d2, r;
        int j;
        entry_geometry(lt, w, pi, i * w.cap + q, d0, d1, d2, r, j);
        R fc, dfc;
        cutoff_fn<R>(m.cutoff_type, (R)r, rc, rrc, fc, dfc);
        dx[q] = (R)d0;
        dy[q] = (R)d1;
        dz[q] = (R)d2;
        rr[q] = (R)r;
        fcs[q] = fc;
        dfs[q] = dfc;
        el[q] = (unsigned char)w.eidx[j];
    }
    __syncthreads();
    const int s0 = par_off[e], ns = par_off[e + 1] - s0;
    const long long row0 = dsf ? dsf_off[i] : 0;
    for (int s = 0; s < ns; ++s) {
        const SfParam p = par[s0 + s];
        const R eta = (R)p.eta, rs = (R)p.rs, lam = (R)p.lambda, zeta = (R)p.zeta, pre = (R)p.pre;
        const double c = m.sf_mean[m.sf_off[e] + s], den = m.sf_den[m.sf_off[e] + s];
        const R iden = R(1.0 / den);
        R val = 0, sx = 0, sy = 0, sz = 0;
        for (int q = lane; q < n; q += 64) {
            R gx = 0, gy = 0, gz = 0;
            const int fq = el[q];
            const R rq = rr[q];
            if (p.type == 2) {
                if (fq == p.f && rq < rc) {
                    // G2 = exp(-eta (r - Rs)^2) fc(r); d/dR_q with d = R_i - R_q
                    const R t = M<R>::exp(-eta * (rq - rs) * (rq - rs));
                    val += t * fcs[q];
                    const R k = t * (R(2) * eta * (rq - rs) * fcs[q] / rq - dfs[q]);
                    gx = k * dx[q];
                    gy = k * dy[q];
                    gz = k * dz[q];
                }
            } else if (p.type == 9) {
                if ((fq == p.f || fq == p.g) && rq < rc) {
                    // G5: no fc(r_jk), no r_jk^2 in the exponential
                    const int want = fq == p.f ? p.g : p.f;
                    const R ax = dx[q], ay = dy[q], az = dz[q], fij = fcs[q], dfij = dfs[q];
                    for (int k = 0; k < n; ++k) {
                        if (k == q || el[k] != want || !(rr[k] < rc)) continue;
                        const R bx = dx[k], by = dy[k], bz = dz[k], rk = rr[k], fk = fcs[k];
                        const R ab = ax * bx + ay * by + az * bz;
                        const R inv = R(1) / (rq * rk);
                        const R cost = ab * inv;
                        R base = R(1) + lam * cost;
                        if (base < R(0)) base = R(0);
                        const R P = powz<R>(base, zeta), dP = zeta * lam * powz<R>(base, zeta - R(1));
                        const R E = M<R>::exp(-eta * (rq * rq + rk * rk));
                        const R F = fij * fk;
                        val += pre * P * E * F;
                        const R a3 = cost / (rq * rq);
                        const R cq = P * E * (-fk * dfij) + P * R(2) * eta * E * F + dP * a3 * E * F;
                        const R ck = -dP * inv * E * F;
                        gx += pre * (cq * ax + ck * bx);
                        gy += pre * (cq * ay + ck * by);
                        gz += pre * (cq * az + ck * bz);
                    }
                }
            } else if ((fq == p.f || fq == p.g) && rq < rc) {
                const int want = fq == p.f ? p.g : p.f;   // the partner's element
                const R ax = dx[q], ay = dy[q], az = dz[q], fij = fcs[q], dfij = dfs[q];
                for (int k = 0; k < n; ++k) {
                    if (k == q || el[k] != want || !(rr[k] < rc)) continue;
                    const R bx = dx[k], by = dy[k], bz = dz[k], rk = rr[k];
                    // d_jk = R_q - R_k = d_ik - d_iq
                    const R cx = bx - ax, cy = by - ay, cz = bz - az;
                    const R rjk2 = cx * cx + cy * cy + cz * cz;
                    if (!(rjk2 < rc * rc)) continue;
                    const R rjk = M<R>::sqrt(rjk2);
                    R fjk, dfjk;
                    cutoff_in<R>(m.cutoff_type, rjk, rrc, fjk, dfjk);
                    const R ab = ax * bx + ay * by + az * bz;
                    const R inv = R(1) / (rq * rk);
                    const R cost = ab * inv;
                    R base = R(1) + lam * cost;
                    if (base < R(0)) base = R(0);
                    const R P = powz<R>(base, zeta);
                    const R dP = zeta * lam * powz<R>(base, zeta - R(1));
                    const R E = M<R>::exp(-eta * (rq * rq + rk * rk + rjk2));
                    const R F = fij * fcs[k] * fjk;
                    val += pre * P * E * F;
                    // d cos / dR_q = -d_ik / (rij rik) + d_ij (d_ij . d_ik) / (rij^3 rik)
                    const R a3 = ab * inv / (rq * rq);
                    const R dcx = -bx * inv + ax * a3, dcy = -by * inv + ay * a3, dcz = -bz * inv + az * a3;
                    // dF / dR_q = fc_ik (fc'_jk / r_jk d_jk fc_ij - fc'_ij / r_ij d_ij fc_jk)
                    const R fk = fcs[k];
                    const R u = fk * dfjk * fij, v = fk * dfij * fjk;
                    const R dFx = u * cx - v * ax, dFy = u * cy - v * ay, dFz = u * cz - v * az;
                    // dE / dR_q = 2 eta E (d_ij - d_jk)
                    const R k2 = R(2) * eta * E;
                    const R dEx = k2 * (ax - cx), dEy = k2 * (ay - cy), dEz = k2 * (az - cz);
                    gx += pre * (P * E * dFx + P * dEx * F + dP * dcx * E * F);
                    gy += pre * (P * E * dFy + P * dEy * F + dP * dcy * E * F);
                    gz += pre * (P * E * dFz + P * dEz * F + dP * dcz * E * F);
                }
            }
            sx += gx;
            sy += gy;
            sz += gz;
            if (dsf) {
                const long long o = row0 + 3 * ((long long)s * (n + 1) + 1 + q);
                dsf[o + 0] = (double)(gx * iden);
                dsf[o + 1] = (double)(gy * iden);
                dsf[o + 2] = (double)(gz * iden);
            }
        }
        val = wave_total(val);
        sx = wave_total(sx);
        sy = wave_total(sy);
        sz = wave_total(sz);
        if (lane == 0) {
            const double G = p.type == 2 ? (double)val : 0.5 * (double)val;   // pairs seen from both ends
            Gout[(size_t)i * ldG + s] = (G - c) / den;
            if (dsf) {
                const long long o = row0 + 3 * ((long long)s * (n + 1));
                dsf[o + 0] = -(double)(sx * iden);
                dsf[o + 1] = -(double)(sy * iden);
                dsf[o + 2] = -(double)(sz * iden);
            }
        }
    }
}

// ---------------------------------------------------------------- G5
// RuNNer's wide angular function (type 9): G5 = 2^(1-zeta) sum_{j<k}
// (1 + lambda cos theta_jik)^zeta exp(-eta (r_ij^2 + r_ik^2)) fc(r_ij) fc(r_ik),
// with no cutoff of r_jk, so the compacted triplet lists of the G4 passes
// (r_jk < rc) do not serve it. One wavefront per atom; the neighbours within
// the cutoff are staged in LDS grouped by element, in list order within an
// element (ballot ranks, so the order never depends on timing), as
// d = R_i - R_n, r, fc, fc'/r. For a function of element pair {f, g} the
// lanes take the entries q of group f and sum over the partners k of group g
// (k != q): each unordered pair once for f != g, twice for f = g.
template <typename R>
struct G5Lds {
    R *dx, *dy, *dz, *r, *fc, *dfc;
    unsigned char *el;
    unsigned short *ent;
    int seg[kMaxEl + 1];
};

template <typename R>
__host__ __device__ __forceinline__ size_t g5_lds_bytes(int lim)
{
    return (size_t)lim * (6 * sizeof(R) + 3) + 16;
}

// stage atom i's entries within rc; returns how many
template <typename R>
__device__ __forceinline__ int g5_stage(const DevModel &m, const Geo &g, const Work &w, int i, R rc, char *base,
                                        G5Lds<R> &L)
{
    const int lane = threadIdx.x, n = w.count[i], lim = w.lim;
    L.dx = (R *)base;
    L.dy = L.dx + lim;
    L.dz = L.dy + lim;
    L.r = L.dz + lim;
    L.fc = L.r + lim;
    L.dfc = L.fc + lim;
    L.el = (unsigned char *)(L.dfc + lim);
    L.ent = (unsigned short *)(((size_t)(L.el + lim) + 1) & ~(size_t)1);
    const double4 pi = w.wpos[i];
    const Lat9 lt = atom_lattice(g, w, i);
    const R rrc = R(1) / rc;
    const unsigned long long below = (1ull << lane) - 1;
    int cnt[kMaxEl];
#pragma unroll
    for (int f = 0; f < kMaxEl; ++f) cnt[f] = 0;
    for (int pass = 0; pass < 2; ++pass) {
        int run[kMaxEl];
        if (pass == 1) {   // group starts
            int o = 0;
#pragma unroll
            for (int f = 0; f < kMaxEl; ++f) {
                L.seg[f] = o;
                run[f] = o;
                o += cnt[f];
            }
            L.seg[kMaxEl] = o;
        }
        for (int c0 = 0; c0 < n; c0 += 64) {
            const int q = c0 + lane;
            int f = -1;
            double d0 = 0, d1 = 0, d2 = 0, r = 0;
            if (q < n) {
                int j;
                entry_geometry(lt, w, pi, i * w.cap + q, d0, d1, d2, r, j);
                if ((R)r < rc) f = w.eidx[j];
            }
#pragma unroll
            for (int h = 0; h < kMaxEl; ++h) {
                if (h >= m.nel) break;
                const unsigned long long mk = __ballot(f == h);
                if (pass == 1 && f == h) {
                    const int at = run[h] + __popcll(mk & below);
                    R fc, dfc;
                    cutoff_in<R>(m.cutoff_type, (R)r, rrc, fc, dfc);
                    L.dx[at] = (R)d0;
                    L.dy[at] = (R)d1;
                    L.dz[at] = (R)d2;
                    L.r[at] = (R)r;
                    L.fc[at] = fc;
                    L.dfc[at] = dfc;
                    L.el[at] = (unsigned char)h;
                    L.ent[at] = (unsigned short)q;
                }
                if (pass == 0) cnt[h] += __popcll(mk);
                else run[h] += __popcll(mk);
            }
        }
    }
    __syncthreads();
    return L.seg[kMaxEl];
}

template <typename R>
__global__ __launch_bounds__(64) void k_g5_values(DevModel m, Geo g, Work w, R *G)
{
    extern __shared__ __align__(16) char smem[];
    const int i = blockIdx.x, lane = threadIdx.x, e = w.eidx[i];
    const int t0 = m.g5_start[e], nt = m.g5_count[e];
    if (nt == 0) return;
    const R rc = (R)m.rc[e];
    __shared__ G5Lds<R> L;
    G5Lds<R> Lr;
    g5_stage<R>(m, g, w, i, rc, smem, Lr);
    if (lane == 0) L = Lr;
    __syncthreads();
    const G5Lds<R> &S = L;
    for (int t = 0; t < nt; ++t) {
        const SfParam p = m.g5[t0 + t];
        const R eta = (R)p.eta, lam = (R)p.lambda, zeta = (R)p.zeta, pre = (R)p.pre;
        const int fa = S.seg[p.f], fb = S.seg[p.f + 1], ga = S.seg[p.g], gb = S.seg[p.g + 1];
        R val = 0;
        for (int q = fa + lane; q < fb; q += 64) {
            const R ax = S.dx[q], ay = S.dy[q], az = S.dz[q], rq = S.r[q], fq = S.fc[q];
            for (int k = ga; k < gb; ++k) {
                if (k == q) continue;
                const R rk = S.r[k];
                const R cost = (ax * S.dx[k] + ay * S.dy[k] + az * S.dz[k]) / (rq * rk);
                R base = R(1) + lam * cost;
                if (base < R(0)) base = R(0);
                val += powz<R>(base, zeta) * M<R>::exp(-eta * (rq * rq + rk * rk)) * fq * S.fc[k];
            }
        }
        val = wave_total(val);
        if (lane == 0) {
            const int s = m.sf_off[e] + p.s;
            const double v = (double)pre * (p.f == p.g ? 0.5 * (double)val : (double)val);
            G[(size_t)i * m.nsfmax + p.s] = (R)((v - m.sf_mean[s]) / m.sf_den[s]);
        }
    }
}

template <typename R>
__global__ __launch_bounds__(64) void k_g5_grad(DevModel m, Geo g, Work w, const R *dEdG, R *gent)
{
    extern __shared__ __align__(16) char smem[];
    const int i = blockIdx.x, lane = threadIdx.x, e = w.eidx[i];
    const int t0 = m.g5_start[e], nt = m.g5_count[e];
    if (nt == 0) return;
    const R rc = (R)m.rc[e];
    __shared__ G5Lds<R> L;
    G5Lds<R> Lr;
    const int n = g5_stage<R>(m, g, w, i, rc, smem, Lr);
    if (lane == 0) L = Lr;
    double *gx = (double *)(smem + ((g5_lds_bytes<R>(w.lim) + 15) / 16 * 16)), *gy = gx + w.lim, *gz = gy + w.lim;
    for (int q = lane; q < n; q += 64) gx[q] = gy[q] = gz[q] = 0.0;
    __syncthreads();
    const G5Lds<R> &S = L;
    for (int t = 0; t < nt; ++t) {
        const SfParam p = m.g5[t0 + t];
        const int s = m.sf_off[e] + p.s;
        const R wgt = (R)((double)dEdG[(size_t)i * m.nsfmax + p.s] / m.sf_den[s]);
        const R eta = (R)p.eta, lam = (R)p.lambda, zeta = (R)p.zeta, pre = (R)p.pre;
        // lanes over group f with partners in g, then (f != g) over g with partners in f
        for (int side = 0; side < (p.f == p.g ? 1 : 2); ++side) {
            const int own = side == 0 ? p.f : p.g, oth = side == 0 ? p.g : p.f;
            const int fa = S.seg[own], fb = S.seg[own + 1], ga = S.seg[oth], gb = S.seg[oth + 1];
            for (int q = fa + lane; q < fb; q += 64) {
                const R ax = S.dx[q], ay = S.dy[q], az = S.dz[q], rq = S.r[q], fq = S.fc[q], dfq = S.dfc[q];
                R sx = 0, sy = 0, sz = 0;
                for (int k = ga; k < gb; ++k) {
                    if (k == q) continue;
                    const R bx = S.dx[k], by = S.dy[k], bz = S.dz[k], rk = S.r[k], fk = S.fc[k];
                    const R ab = ax * bx + ay * by + az * bz;
                    const R inv = R(1) / (rq * rk);
                    const R cost = ab * inv;
                    R base = R(1) + lam * cost;
                    if (base < R(0)) base = R(0);
                    const R P = powz<R>(base, zeta), dP = zeta * lam * powz<R>(base, zeta - R(1));
                    const R E = M<R>::exp(-eta * (rq * rq + rk * rk));
                    const R F = fq * fk;
                    // d cos/dR_q = -d_k/(r_q r_k) + d_q cos / r_q^2; dF/dR_q = -fc_k fc'_q/r_q d_q;
                    // dE/dR_q = 2 eta E d_q (d = R_i - R_n)
                    const R a3 = cost / (rq * rq);
                    const R cq = P * E * (-fk * dfq) + P * R(2) * eta * E * F + dP * a3 * E * F;
                    const R ck = -dP * inv * E * F;
                    sx += cq * ax + ck * bx;
                    sy += cq * ay + ck * by;
                    sz += cq * az + ck * bz;
                }
                gx[q] += (double)(wgt * pre * sx);
                gy[q] += (double)(wgt * pre * sy);
                gz[q] += (double)(wgt * pre * sz);
            }
            __syncthreads();
        }
    }
    // the rows of the entries, and the atom's virial -sum_n g_n (x) d_in
    double v[9];
#pragma unroll
    for (int t = 0; t < 9; ++t) v[t] = 0;
    for (int q = lane; q < n; q += 64) {
        const size_t en = (size_t)i * w.cap + S.ent[q];
        gent[3 * en + 0] += (R)gx[q];
        gent[3 * en + 1] += (R)gy[q];
        gent[3 * en + 2] += (R)gz[q];
        const double dxq = S.dx[q], dyq = S.dy[q], dzq = S.dz[q];
        v[0] -= gx[q] * dxq; v[1] -= gx[q] * dyq; v[2] -= gx[q] * dzq;
        v[3] -= gy[q] * dxq; v[4] -= gy[q] * dyq; v[5] -= gy[q] * dzq;
        v[6] -= gz[q] * dxq; v[7] -= gz[q] * dyq; v[8] -= gz[q] * dzq;
    }
#pragma unroll
    for (int t = 0; t < 9; ++t) {
        const double sv = wave_total(v[t]);
        if (lane == 0) w.vir[9 * (size_t)i + t] += sv;
    }
}

template <typename R>
void launch_g5_values(const DevModel &m, const Geo &g, const Work &w, R *G, hipStream_t s)
{
    if (w.nat == 0 || m.g5_total == 0) return;
    hipLaunchKernelGGL(k_g5_values<R>, dim3(w.nat), dim3(64), g5_lds_bytes<R>(w.lim), s, m, g, w, G);
}
template <typename R>
void launch_g5_grad(const DevModel &m, const Geo &g, const Work &w, const R *dEdG, R *gent, hipStream_t s)
{
    if (w.nat == 0 || m.g5_total == 0) return;
    const size_t lds = (g5_lds_bytes<R>(w.lim) + 15) / 16 * 16 + 24 * (size_t)w.lim;
    hipLaunchKernelGGL(k_g5_grad<R>, dim3(w.nat), dim3(64), lds, s, m, g, w, dEdG, gent);
}
template void launch_g5_values<double>(const DevModel &, const Geo &, const Work &, double *, hipStream_t);
template void launch_g5_values<float>(const DevModel &, const Geo &, const Work &, float *, hipStream_t);
template void launch_g5_grad<double>(const DevModel &, const Geo &, const Work &, const double *, double *, hipStream_t);
template void launch_g5_grad<float>(const DevModel &, const Geo &, const Work &, const float *, float *, hipStream_t);

template <typename R>
void launch_sf_single(const DevModel &m, const Geo &g, const Work &w, const SfParam *par, const int *par_off,
                      const long long *dsf_off, int ldG, double *G, double *dsf, hipStream_t s)
{
    if (w.nat == 0) return;
    const size_t lds = (size_t)w.lim * (6 * sizeof(R) + 1) + 16;
    hipLaunchKernelGGL(k_sf_single<R>, dim3(w.nat), dim3(64), lds, s, m, g, w, par, par_off, dsf_off, ldG, G, dsf);
}
template void launch_sf_single<double>(const DevModel &, const Geo &, const Work &, const SfParam *, const int *,
                                       const long long *, int, double *, double *, hipStream_t);
template void launch_sf_single<float>(const DevModel &, const Geo &, const Work &, const SfParam *, const int *,
                                      const long long *, int, double *, double *, hipStream_t);

// LDS of one wavefront of the kernels serving nmax neighbours with one atom
// per wavefront (this decides the staging capacity limit)
size_t sf_lds_bytes(bool fp64, int nmax, const DevModel &m, bool grad)
{
    nmax = sf_bucket(nmax);
    const int nsfmax = m.nsfmax, g4max = m.g4_max, nel = m.nel;
    if (!fp64 && !grad && sf_direct(m, 1, nmax)) return direct_offsets(nmax, nsfmax, g4max, m.g4_nblk).total;
    if (!grad && sf_allpairs_enabled())
        return fp64 ? lds_bytes<double>(nmax, 1, nsfmax, g4max, nel, false, true)
                    : lds_bytes<float>(nmax, 1, nsfmax, g4max, nel, false, true);
    return fp64 ? lds_bytes<double>(nmax, 1, nsfmax, g4max, nel, grad)
                : lds_bytes<float>(nmax, 1, nsfmax, g4max, nel, grad);
}

// the tile the kernels use for a requested one: ga when that variant is built
// for the staging capacity and its wavefront fits 64 KiB of LDS in both
// passes, else the largest smaller one that does (1 always does)
int sf_tile(bool fp64, int nmax, int ga, int nsfmax, int g4max, int nel)
{
    // tiles of 2 and 4 atoms are built for staging sizes 64 and 128
    const int b = sf_bucket(nmax) <= 64 ? 64 : sf_bucket(nmax) <= 128 ? 128 : 0;
    for (int a = ga >= 4 ? 4 : (ga >= 2 ? 2 : 1); a > 1; a /= 2) {
        if (b == 0) continue;
        const size_t v = fp64 ? lds_bytes<double>(b, a, nsfmax, g4max, nel, false)
                              : lds_bytes<float>(b, a, nsfmax, g4max, nel, false);
        const size_t gr = fp64 ? lds_bytes<double>(b, a, nsfmax, g4max, nel, true)
                               : lds_bytes<float>(b, a, nsfmax, g4max, nel, true);
        if (v <= 65536 && gr <= 65536) return a;
    }
    return 1;
}

}  // namespace hdnnp
