// Wave-per-centre symmetry-function kernels for gfx950.
//
// One 64-lane wavefront owns one centre atom at a time; a persistent grid
// walks the centres, each XCD one contiguous slice of them.  Per centre:
//   1. pair phase: lane s takes neighbour-list entry s (at most 64), forms the
//      FP64 displacement and distance, and the entries inside the cutoff are
//      compacted (list order kept) into LDS arrays; each lane adds its g2
//      rows into FP64 accumulators.
//   2. triplet enumeration: the lanes walk all pairs (a < b) of compacted
//      neighbours in rotational order (pair_of) and compact the pairs
//      with r_jk < rc_jk into one LDS region per species-pair block.
//   3. triplet phase: the lanes walk the active pairs only.  cos(theta),
//      fc_jk, one exp per distinct g4 eta and the zeta powers of each lambda's
//      base serve all rows of a triplet; each row adds one FP64 fma.
//   4. the accumulators are summed over the lanes with a fixed-order
//      reduce-scatter butterfly (permlane swaps and DPP) and written into the
//      descriptor row (the reference's += semantics, or = for the serving
//      loop).
// Global loads run one centre ahead of their use: the next centre's record at
// the start of a centre, its neighbour indices after the pair phase, its
// neighbours' species and coordinates after the enumeration, so the chain
// offsets -> neighbour list -> coordinates is hidden behind the triplet work.
// The adjoint kernel forms dE/d r_ij, dE/d r_ik and dE/d r_jk of a triplet
// from three row sums and adds the resulting vectors into per-neighbour FP64
// LDS accumulators (ds_add_f64; the wave owns its centre, so the order is
// fixed); it then writes one record per neighbour-list edge and one for the
// centre, at the positions the atomic-free gather in sf_plan.hip reads.
// Nothing is atomic in global memory, and a centre's results do not depend
// on which other centres share the launch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "sf_device.h"
#include "sf_kernels.h"

namespace ldsf {

namespace {

constexpr int kW = 64;

template <int K>
struct Log2;
template <> struct Log2<1> { static constexpr int v = 0; };
template <> struct Log2<2> { static constexpr int v = 1; };
template <> struct Log2<4> { static constexpr int v = 2; };
template <> struct Log2<8> { static constexpr int v = 3; };
template <> struct Log2<16> { static constexpr int v = 4; };

// bits of the wave mask m set in the lanes below this one (v_mbcnt: the lane
// is implicit; every caller's s is its lane)
__device__ __forceinline__ int below(unsigned long long m, int)
{
    return (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                          __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

// Cross-lane exchanges of the reduction butterfly, for partner bit M.
// M = 32, 16: v_permlane32_swap / v_permlane16_swap exchange the two halves
// of a pair of registers between lane halves (rows), so that after
// xswap<M>(a, b) every lane's a + b is its own kept value plus the partner's
// (lanes with bit M clear keep a, lanes with it set keep b).
template <int M>
__device__ __forceinline__ void xswap(double &a, double &b)
{
    static_assert(M == 32 || M == 16, "permlane swaps exchange 32- or 16-lane halves");
    const unsigned long long ua = __double_as_longlong(a), ub = __double_as_longlong(b);
    const unsigned al = (unsigned)ua, ah = (unsigned)(ua >> 32);
    const unsigned bl = (unsigned)ub, bh = (unsigned)(ub >> 32);
    unsigned l0, l1, h0, h1;
    if constexpr (M == 32) {
        const auto rl = __builtin_amdgcn_permlane32_swap(al, bl, false, false);
        const auto rh = __builtin_amdgcn_permlane32_swap(ah, bh, false, false);
        l0 = rl[0]; l1 = rl[1]; h0 = rh[0]; h1 = rh[1];
    } else {
        const auto rl = __builtin_amdgcn_permlane16_swap(al, bl, false, false);
        const auto rh = __builtin_amdgcn_permlane16_swap(ah, bh, false, false);
        l0 = rl[0]; l1 = rl[1]; h0 = rh[0]; h1 = rh[1];
    }
    a = __longlong_as_double((long long)(((unsigned long long)h0 << 32) | l0));
    b = __longlong_as_double((long long)(((unsigned long long)h1 << 32) | l1));
}

// M <= 8: the partner lane's value through DPP: lane ^ 1 and lane ^ 2 (quad
// permutations), the mirror lane in its 8-lane half row (differs in bit 2)
// and in its 16-lane row (differs in bit 3).  (These DPP calls, and the row
// rotations below, take 0 for a source lane they cannot read, whether as the
// old value or, with bound_ctrl set, as the bound value; bound_ctrl spares the
// register copy of the old value.)
template <int M>
__device__ __forceinline__ double dpp_partner(double v)
{
    constexpr int ctrl = M == 1 ? 0xB1 : M == 2 ? 0x4E : M == 4 ? 0x141 : 0x140;
    static_assert(M == 1 || M == 2 || M == 4 || M == 8, "DPP partner within a row");
    const unsigned long long u = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)u, ctrl, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(u >> 32), ctrl, 0xF, 0xF, true);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) |
                                            (unsigned)lo));
}

// Recursive-halving reduce-scatter of K values over the 64 lanes: step with H
// live values and partner bit M, all compile-time.  Each step pairs lanes that
// differ in bit M and agree in the bits already reduced, so every lane's
// partial sums stay disjoint.
template <int K, int H, int M>
struct Halve {
    static __device__ __forceinline__ void run(double (&v)[K], int s)
    {
        if constexpr (M >= 16) {
#pragma unroll
            for (int q = 0; q < H; ++q) {
                double a = v[q], b = v[q + H];
                xswap<M>(a, b);
                v[q] = a + b;
            }
        } else {
            const bool hi = (s & M) != 0;
#pragma unroll
            for (int q = 0; q < H; ++q) {
                const double send = hi ? v[q] : v[q + H];
                const double keep = hi ? v[q + H] : v[q];
                v[q] = keep + dpp_partner<M>(send);
            }
        }
        Halve<K, H / 2, M / 2>::run(v, s);
    }
};

template <int K, int M>
struct Halve<K, 0, M> {
    static __device__ __forceinline__ void run(double (&v)[K], int)
    {
        if constexpr (M >= 16) {
            double a = v[0], b = v[0];
            xswap<M>(a, b);
            v[0] = a + b;
        } else {
            v[0] += dpp_partner<M>(v[0]);
        }
        Halve<K, 0, M / 2>::run(v, 0);
    }
};

template <int K>
struct Halve<K, 0, 0> {
    static __device__ __forceinline__ void run(double (&)[K], int) {}
};

// Afterwards v[0] of lane s holds the wave total of value (s >> (6 - log2 K));
// the butterfly order is fixed.
template <int K>
__device__ __forceinline__ void wave_reduce(double (&v)[K], int s)
{
    Halve<K, K / 2, kW / 2>::run(v, s);
}

// ceil(c / d) for wave-uniform 0 <= c + d - 1 < 4096 and 1 <= d <= 64 without
// a division: floor(x m_d / 2^20) = floor(x / d) there, m_d = ceil(2^20 / d)
// (checked exhaustively over that range; x m_d < 2^32)
struct Rcp20 {
    unsigned m[65];
};

constexpr Rcp20 make_rcp20()
{
    Rcp20 t{};
    for (int d = 1; d <= 64; ++d) t.m[d] = (1u << 20) / d + ((1u << 20) % d != 0 ? 1u : 0u);
    return t;
}

__device__ const Rcp20 kRcp20 = make_rcp20();

__device__ __forceinline__ int ceil_div_small(int c, int d)
{
    return (int)(((unsigned)(c + d - 1) * kRcp20.m[d]) >> 20);
}

// The g2 rows' sums over a centre's compacted neighbours, in FP64: terms
// holds 8 terms per neighbour (row q of neighbour a at 8 a + q, FP32 or FP64).
// Lane s adds, in list order, row q = s & 7 over one part of its block:
// with two species blocks (neighbours [0, n0) and [n0, M), species-sorted)
// block (s >> 3) & 1 in 4 parts by s >> 4; with one block all M neighbours in
// 8 parts by s >> 3.  The parts are then added across lanes (lane bits 3, 4
// and 5) in a fixed order, every lane of a group ending with the same total,
// and the first lane of each (block, row) adds it into the centre's row.
template <int NSB, typename T>
__device__ __forceinline__ void g2_rows_sum(const T *terms, int M, int n0, int s, const int *slots,
                                            double *lrow)
{
    const int q = s & 7;
    int lo = 0, hi = M, part;
    if constexpr (NSB == 2) {
        const bool b1 = (s & 8) != 0;
        lo = b1 ? n0 : 0;
        hi = b1 ? M : n0;
        part = s >> 4;
    } else {
        part = s >> 3;
    }
    constexpr int kParts = NSB == 2 ? 4 : 8;
    const int len = hi - lo;
    const int a0 = lo + ((len * part) / kParts), a1 = lo + ((len * (part + 1)) / kParts);
    // four loads in flight per step; the padding adds +0.0, which leaves the
    // sum of the non-negative terms bitwise unchanged
    double sum = 0.0;
    for (int a = a0; a < a1; a += 4) {
        T t[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) t[k] = a + k < a1 ? terms[8 * (a + k) + q] : (T)0;
#pragma unroll
        for (int k = 0; k < 4; ++k) sum += (double)t[k];
    }
    if constexpr (NSB == 1) {
        // lane ^ 8 within its 16-lane row: DPP row_ror:8
        const unsigned long long u = __double_as_longlong(sum);
        const int lo32 = __builtin_amdgcn_update_dpp(0, (int)(unsigned)u, 0x128, 0xF, 0xF, true);
        const int hi32 = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(u >> 32), 0x128, 0xF, 0xF, true);
        sum += __longlong_as_double((long long)(((unsigned long long)(unsigned)hi32 << 32) |
                                                (unsigned)lo32));
    }
    {
        double a = sum, b = sum;
        xswap<16>(a, b);
        sum = a + b;
    }
    {
        double a = sum, b = sum;
        xswap<32>(a, b);
        sum = a + b;
    }
    if (s < (NSB == 2 ? 16 : 8)) {
        const int slot = slots[NSB == 2 ? (s & 15) : q];
        if (slot >= 0) lrow[slot] += sum;
    }
}

// The same sums over 32 lanes, for a centre of at most 32 compacted
// neighbours: each half-wave serves one centre (its terms at `terms`, M
// neighbours, n0 of species 0, descriptor row `lrow`; on == false: a half
// without a centre), lane s & 31 adding row s & 7 over one part of its block
// (2 parts with two blocks, 4 with one), the parts then added across lane
// bits 3 and 4 only.  A centre's sums are the same in either half.
template <int NSB, typename T>
__device__ __forceinline__ void g2_rows_sum_half(const T *terms, int M, int n0, int s,
                                                 const int *slots, double *lrow, bool on)
{
    const int ls = s & 31, q = ls & 7;
    int lo = 0, hi = M, part;
    if constexpr (NSB == 2) {
        const bool b1 = (ls & 8) != 0;
        lo = b1 ? n0 : 0;
        hi = b1 ? M : n0;
        part = ls >> 4;
    } else {
        part = ls >> 3;
    }
    constexpr int kParts = NSB == 2 ? 2 : 4;
    const int len = hi - lo;
    const int a0 = lo + ((len * part) / kParts), a1 = lo + ((len * (part + 1)) / kParts);
    double sum = 0.0;
    for (int a = a0; a < a1; a += 4) {
        T t[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) t[k] = a + k < a1 ? terms[8 * (a + k) + q] : (T)0;
#pragma unroll
        for (int k = 0; k < 4; ++k) sum += (double)t[k];
    }
    if constexpr (NSB == 1) {
        // lane ^ 8 within its 16-lane row: DPP row_ror:8
        const unsigned long long u = __double_as_longlong(sum);
        const int lo32 = __builtin_amdgcn_update_dpp(0, (int)(unsigned)u, 0x128, 0xF, 0xF, true);
        const int hi32 = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(u >> 32), 0x128, 0xF, 0xF, true);
        sum += __longlong_as_double((long long)(((unsigned long long)(unsigned)hi32 << 32) |
                                                (unsigned)lo32));
    }
    {
        double a = sum, b = sum;
        xswap<16>(a, b);
        sum = a + b;
    }
    if (on && ls < (NSB == 2 ? 16 : 8)) {
        const int slot = slots[NSB == 2 ? (ls & 15) : q];
        if (slot >= 0) lrow[slot] += sum;
    }
}

// species-pair index a * ns + b (ns is small: v_mul_u32_u24, or a shift
// where ns is a compile-time 2)
__device__ __forceinline__ int pidx(int a, int b, int ns)
{
    return (int)__umul24(a, ns) + b;
}

// The pairs of M compacted neighbours (M = 2..64) in rotational order: the
// T = M(M-1)/2 candidates t run in rows of M, row r pairing neighbour i with
// (i + r + 1) mod M (for even M the last row, r = M/2 - 1, holds only its M/2
// pairs i < M/2, which end at T): every unordered pair once.  Along a row the
// lanes of one instruction touch distinct neighbours: few LDS bank conflicts,
// and the adjoint's LDS accumulations into a and b rarely collide.  A lane
// computes its pair from t, with no table load (whose latency a pass of this
// length cannot hide): r = floor(t / M) from the float (t + 0.5) / M, which
// lies at least 0.5 / M from an integer, far beyond the rounding of the
// product and of inv_m = 1 / M (within an ulp).  ORDERED: a < b, as the list
// order of the single-block kernels needs for an asymmetric cutoff matrix
// (rcut(j, k) with j first, src/SymmetryFunctions.cpp:206); the species-
// resolved blocks have a symmetric one, and a triplet's terms are symmetric
// in (a, b) up to exactly commuted operations.
template <bool ORDERED>
__device__ __forceinline__ void pair_of(int t, int M, float inv_m, int &a, int &b)
{
    const int r = (int)fmaf((float)t, inv_m, 0.5f * inv_m);
    const int i = t - (int)__umul24(r, M);
    const unsigned c = (unsigned)(i + r + 1);
    const int j = (int)min(c, c - (unsigned)M);  // (i + r + 1) mod M
    a = ORDERED ? min(i, j) : i;
    b = ORDERED ? max(i, j) : j;
}

// centres of this block: blocks are dealt to the 8 XCDs round-robin (block b
// on XCD b % 8); each XCD takes one contiguous slice of the centres
struct TileRange {
    int hi, step, first;
};

__device__ __forceinline__ TileRange tile_range(int n)
{
    const int G = gridDim.x, b = blockIdx.x;
    TileRange r;
    if (G >= 8 && (G & 7) == 0) {
        const int xcd = b & 7, local = b >> 3, per = G >> 3;
        const int chunk = (n + 7) >> 3;
        const int lo = min(n, xcd * chunk);
        r.hi = min(n, lo + chunk);
        r.first = lo + local;
        r.step = per;
    } else {
        r.hi = n;
        r.first = b;
        r.step = G;
    }
    return r;
}

// B^(2^k), k = 0..nsq (nsq uniform): every integer zeta <= 63 of the grid is a
// product of these
template <typename T>
__device__ __forceinline__ void squarings(T b, int nsq, T (&sq)[6])
{
    sq[0] = b;
#pragma unroll
    for (int k = 1; k < 6; ++k) sq[k] = (k <= nsq) ? sq[k - 1] * sq[k - 1] : (T)0;
}

template <typename T>
__device__ __forceinline__ T pow_bits(const T (&sq)[6], int bits)
{
    T r = (T)1;
#pragma unroll
    for (int k = 0; k < 6; ++k)
        if (bits & (1 << k)) r *= sq[k];
    return r;
}

// b^zeta_z (b^(zeta_z - 1) when minus_one) for b > 0, and 0 for b <= 0 as the
// reference's guard (src/SymmetryFunctions.cpp:121)
template <typename T>
__device__ __forceinline__ T power_z(T b, const T (&sq)[6], int z, const WaveParams &wp,
                                     bool minus_one)
{
    const int iz = __builtin_amdgcn_readfirstlane(wp.izeta[z]);
    T v;
    if (iz >= (minus_one ? 1 : 0)) {
        v = pow_bits(sq, minus_one ? iz - 1 : iz);
    } else {
        const double zz = wp.zeta[z] - (minus_one ? 1.0 : 0.0);
        if (sizeof(T) == 8) v = (T)pow((double)b, zz);
        else v = (T)__builtin_amdgcn_exp2f((float)zz * __builtin_amdgcn_logf((float)b));
    }
    return b > (T)0 ? v : (T)0;
}

// B^zeta_z for the nz distinct zetas of the grid (0 for B <= 0).  FP64: the
// squaring table (exact products for integer zeta).  FP32-bulk:
// exp2(zeta log2 B), one v_log_f32 per base and one v_exp_f32 per zeta; the
// relative error, ~zeta |log2 B| 2^-24, is below 1e-6 wherever the term is not
// negligible against the block (B >= 2^-2 for zeta = 16), like the zeta-fold
// growth of B's own FP32 rounding that any evaluation carries.
template <int P>
__device__ __forceinline__ void powers(typename Real<P>::T b, int nsq, const WaveParams &wp,
                                       int nz, typename Real<P>::T (&pw)[4])
{
    using T = typename Real<P>::T;
    if constexpr (P == 0) {
        T sq[6];
        squarings(b, nsq, sq);
#pragma unroll
        for (int z = 0; z < 4; ++z)
            if (z < nz) pw[z] = power_z<T>(b, sq, z, wp, false);
    } else {
        const float lb = __builtin_amdgcn_logf(b);
#pragma unroll
        for (int z = 0; z < 4; ++z)
            pw[z] = b > 0.0f ? __builtin_amdgcn_exp2f((float)wp.zeta[z] * lb) : 0.0f;
    }
}

// B^(zeta_z - 1), the same way
template <int P>
__device__ __forceinline__ void powers_m1(typename Real<P>::T b, int nsq, const WaveParams &wp,
                                          int nz, typename Real<P>::T (&pm)[4])
{
    using T = typename Real<P>::T;
    if constexpr (P == 0) {
        T sq[6];
        squarings(b, nsq, sq);
#pragma unroll
        for (int z = 0; z < 4; ++z)
            if (z < nz) pm[z] = power_z<T>(b, sq, z, wp, true);
    } else {
        const float lb = __builtin_amdgcn_logf(b);
#pragma unroll
        for (int z = 0; z < 4; ++z)
            pm[z] = b > 0.0f ? __builtin_amdgcn_exp2f(((float)wp.zeta[z] - 1.0f) * lb) : 0.0f;
    }
}

// ---- LDS ----

// per-wave tables: staged parameters and species-pair cutoff data
struct Shared {
    WaveParams *par;
    double *rc, *rcinv, *rc2, *rcs;  // rc, 1/rc, rc^2, (pi/2)/rc per species pair (ns^2 each)
    double *rct;                     // the least x with sqrt(x) >= rc (FP64, correctly rounded)
};

// the centre's arrays: 64 entries per neighbour array, then the pair list
struct Grp {
    double *px, *py, *pz, *r2, *r, *rinv, *fc;  // absolute neighbour coordinates (r2: unused slot)
    double *dfc, *fx, *fy, *fz, *w, *wv, *g;    // adjoint only (w, wv: row weights of up to
                                                // 3 blocks; g: the centre's dE/dzeta row)
    double *d;                                  // forward only: the centre's descriptor row
    double *gB, *dB;                            // paired: the second centre's g and d rows
    int *sp, *ip;                               // species; adjoint: record position of the edge
    unsigned short *l0;                         // active pairs, one region per species-pair block
};

__host__ __device__ constexpr size_t shared_doubles(int ns)
{
    // an even count: the neighbour arrays that follow hold 16-byte vectors
    return ((sizeof(WaveParams) + 15) / 16) * 2 + 6 * (size_t)ns * ns;
}

// The per-neighbour arrays sit at fixed offsets (64 entries each) from one
// base, so one address per neighbour index serves every array through the
// LDS instructions' immediate offsets; the pair list (tmax entries) follows.
// (stride 65, not 64: arrays read or updated together fall in different banks)
// The adjoint kinds put the row weights (w, wv), the centre's dE/dzeta row (g)
// and, fused, its descriptor row (d) first, so that the triplet loop reaches
// every weight through a ds_read2 immediate offset (at most 2040 bytes) from
// one per-lane address.  The fused kind holds two centres at a time: the
// second centre's neighbours at slots 32..63, its dE/dzeta and descriptor
// rows (gB, dB) after the neighbour arrays.
// The fused kind's three pair blocks' 16 row weights sit kWBlk elements
// apart: its columns of lanes are dealt to the blocks across each 16-lane row,
// so the lanes of one LDS read serve several blocks, each reading the same
// element of its own block, and with a stride of 16 elements those addresses
// would share banks (FP32: blocks 0 and 2; FP64: all three).  With 20 they
// fall on banks 0, 20, 8 (FP32) and 0, 8, 16 (FP64); 80 and 160 bytes keep
// each block 16-byte aligned.  The adjoint kind deals runs of consecutive
// lanes to the blocks (only the 16-lane rows at a run's end mix two blocks)
// and keeps them 16 apart.
constexpr int kWBlk = 20;
constexpr int kWEnd = 2 * kWBlk + 16;                          // doubles of w and of wv (fused)
constexpr int kWBlkAdj = 16;
constexpr int kWEndAdj = 3 * kWBlkAdj;                         // (adjoint)
constexpr int kStr = kW + 1;
constexpr int kHeadAdj = kWEndAdj + kWEndAdj + kWaveSlots;     // w, wv, g
constexpr int kHeadPair = kWEnd + kWEnd + 2 * kWaveSlots;      // w, wv, g, d
constexpr int kFwdDoubles = 8 * kStr;                          // px..fc, d
constexpr int kAdjDoubles = kHeadAdj + 11 * kStr + 1;          // w, wv, g, px..fz, pad
constexpr int kPairRows = kHeadPair + 11 * kStr + 1;           // w, wv, g, d, px..fz, pad
constexpr int kPairDoubles = kPairRows + 2 * kWaveSlots;       // ..., gB, dB
// The FP32-bulk adjoint kinds keep, in place of px..fc, per compacted
// neighbour: D = two double2 (d_x, d_y) and (d_z, r); V = float4 (1/r, fc,
// dfc, and the species, or in the adjoint kind with two blocks, whose sorted
// slots give the species, r in FP32); W = float2 (the two exp(-eta_e r^2) of
// the triplet sets).

// kernel kinds: the forward, the adjoint, the fused forward + adjoint
enum { kFwd = 0, kAdj = 1, kPair = 2 };

__host__ __device__ constexpr int kind_doubles(int kind)
{
    return kind == kFwd ? kFwdDoubles : (kind == kAdj ? kAdjDoubles : kPairDoubles);
}

// The column sums of the triplet passes: 16 values of 16 columns, value k of
// column c at kColStr k + c.  The stores (16 lanes: one value, 16 columns)
// and the row sums (16 lanes: 16 values, one column) each fall on 16
// different bank pairs with a stride of 17 doubles; 16 would put a row sum's
// 16 reads on one bank.
constexpr int kColStr = 17;

// The paired kind's pair-list region also holds the g2 terms of both centres
// (slots 0..63) and, after a centre's triplet passes, its column sums.
constexpr size_t kPairListMin = 16 * kColStr * sizeof(double);

// The pair list's region also holds, between the pair phase and the pair
// enumeration, the g2 terms of every compacted neighbour (8 per neighbour,
// `term_bytes` = the compute type's size each) for the kinds that form the
// descriptors.
__host__ __device__ inline size_t group_bytes(int maxn, int kind, int term_bytes)
{
    const int pairs = maxn * (maxn - 1) / 2;
    size_t list = sizeof(unsigned short) * (size_t)(pairs > 1 ? pairs : 1);
    const size_t terms = (size_t)8 * term_bytes * (kind == kPair ? kW : maxn);
    if (kind != kAdj && terms > list) list = terms;
    if (kind == kPair && list < kPairListMin) list = kPairListMin;
    size_t b = sizeof(double) * kind_doubles(kind) + sizeof(int) * (kind == kFwd ? 1 : 2) * kW + list;
    return (b + 15) & ~size_t(15);
}

__device__ __forceinline__ Shared carve_shared(double *base, int ns)
{
    Shared S;
    S.par = reinterpret_cast<WaveParams *>(base);
    double *p = base + (sizeof(WaveParams) + 7) / 8;
    S.rc = p; p += ns * ns;
    S.rcinv = p; p += ns * ns;
    S.rc2 = p; p += ns * ns;
    S.rcs = p; p += ns * ns;
    S.rct = p;
    return S;
}

__device__ __forceinline__ Grp carve_group(double *base, int ns, int kind)
{
    const bool adjoint = kind != kFwd;
    double *p = base + shared_doubles(ns);  // 16-byte aligned
    Grp L;
    double *q = p;
    if (adjoint) {
        const int wend = kind == kPair ? kWEnd : kWEndAdj;
        L.w = p;  // read as float4 / double2
        L.wv = p + wend;
        L.g = p + 2 * wend;
        L.d = kind == kPair ? p + 2 * wend + kWaveSlots : nullptr;
        q = p + (kind == kPair ? kHeadPair : kHeadAdj);  // even: 16-byte aligned
    }
    L.gB = kind == kPair ? p + kPairRows : nullptr;
    L.dB = kind == kPair ? p + kPairRows + kWaveSlots : nullptr;
    L.px = q;
    L.py = q + kStr;
    L.pz = q + 2 * kStr;
    L.r2 = q + 3 * kStr;
    L.r = q + 4 * kStr;
    L.rinv = q + 5 * kStr;
    L.fc = q + 6 * kStr;
    if (adjoint) {
        L.dfc = q + 7 * kStr;
        L.fx = q + 8 * kStr;
        L.fy = q + 9 * kStr;
        L.fz = q + 10 * kStr;
    } else {
        L.dfc = L.fx = L.fy = L.fz = L.w = L.wv = L.g = nullptr;
        L.d = q + 7 * kStr;
    }
    int *ip = reinterpret_cast<int *>(p + kind_doubles(kind));
    L.sp = ip;
    L.ip = adjoint ? ip + kW : nullptr;
    L.l0 = reinterpret_cast<unsigned short *>(ip + (adjoint ? 2 : 1) * kW);
    return L;
}

// staging shared by both kernels; ends with a barrier
__device__ __forceinline__ void stage_shared(const Shared &S, const WaveTab &wt, int lane)
{
    const int ns = wt.n_species;
    for (int k = lane; k < ns * ns; k += kW) {
        const double rc = wt.rc[k];
        S.rc[k] = rc;
        S.rcinv[k] = 1.0 / rc;
        S.rc2[k] = rc * rc;
        S.rcs[k] = 0.5 * kPi / rc;
        // sqrt(x) < rc exactly when x < rct: FP64 sqrt is correctly rounded
        // and so monotone, and rct is the least double whose root reaches rc
        double y = rc * rc;
        while (y > 0.0 && sqrt(y) >= rc) y = nextafter(y, 0.0);
        while (sqrt(y) < rc) y = nextafter(y, __builtin_inf());
        S.rct[k] = y;
    }
    const int *src = reinterpret_cast<const int *>(wt.par);
    int *dst = reinterpret_cast<int *>(S.par);
    for (int k = lane; k < (int)(sizeof(WaveParams) / 4); k += kW) dst[k] = src[k];
    __syncthreads();
}

// ---- loads one centre ahead ----

// a centre's record (stage 1)
struct Ctr {
    int c, beg, N, si;
    bool live, work;  // live: c < n_centers; work: live and its request not flagged
    double x0, x1, x2;
};

__device__ __forceinline__ Ctr load_centre(int c, int n_centers, const int *centers,
                                           const int *offs, const int *species, const double *x,
                                           const DevCounts &dc)
{
    Ctr C;
    C.c = c;
    C.live = c < n_centers;
    C.work = false;
    C.beg = C.N = C.si = 0;
    C.x0 = C.x1 = C.x2 = 0.0;
    if (C.live && !(dc.req_flag && dc.req_flag[dc.req_of[c]])) {
        C.work = true;
        const int i = centers ? centers[c] : c;
        C.beg = offs[c];
        C.N = offs[c + 1] - C.beg;
        C.si = species[i];
        C.x0 = x[3 * (size_t)i];
        C.x1 = x[3 * (size_t)i + 1];
        C.x2 = x[3 * (size_t)i + 2];
    }
    return C;
}

// lane s's neighbour of a centre: its index (stage 2), then species and
// absolute coordinates (stage 3)
struct Nbr {
    int sj;
    double p0, p1, p2;
};

__device__ __forceinline__ int load_index(const Ctr &C, int s, const int *nl)
{
    return s < C.N ? nl[C.beg + s] : -1;
}

// (packed: the entry j carries the neighbour's species code above its row,
// a plan's list; else the species array is read)
__device__ __forceinline__ Nbr load_nbr(int j, const int *species, const double *x, bool packed)
{
    Nbr B;
    B.sj = 0;
    B.p0 = B.p1 = B.p2 = 0.0;
    if (j >= 0) {
        if (packed) {
            B.sj = j >> kNlSpeciesShift;
            j &= kNlIndexMask;
        } else {
            B.sj = species[j];
        }
        B.p0 = x[3 * (size_t)j];
        B.p1 = x[3 * (size_t)j + 1];
        B.p2 = x[3 * (size_t)j + 2];
    }
    return B;
}

// cutoff value and derivative of a neighbour at distance r < rc; in the
// FP32-bulk form, y = (pi/2)(rc - r)/rc is formed from the exact FP64 rc - r
__device__ __forceinline__ void cut_pair(double r, double rc, double rcinv, double rcs, double &fc,
                                         double &dfc, Real<0>)
{
    double s, c;
    sincos(ref_arg(r, rc), &s, &c);
    fc = 0.5 * (c + 1.0);
    dfc = -0.5 * s * (kPi * rcinv);
}

__device__ __forceinline__ void cut_pair(double r, double rc, double rcinv, double rcs, float &fc,
                                         float &dfc, Real<1>)
{
    float s, c;
    sincos_quarter((float)((rc - r) * rcs), s, c);
    fc = s * s;
    dfc = -(float)(2.0 * rcs) * s * c;
}

// ---- the FP32-bulk triplet path ----
//
// Per compacted neighbour a the pair phase stores, in the space of the FP64
// neighbour arrays (which this form does not use):
//   D[a]       = (x_a - x_i, y_a - y_i)              double2
//   D[64 + a]  = (z_a - z_i, r_a)                     double2
//   V[a]       forward: (sqrt(1/2) / r_a, g_0, g_1, species bits)   float4
//              with g_e = fc(r_a) exp(-eta_e r_a^2)
//              adjoint: (1 / r_a, fc(r_a), dfc/dr(r_a), species bits)
//   W[a]       adjoint: (exp(-eta_0 r_a^2), exp(-eta_1 r_a^2))     float2
// r_jk^2, d_a . d_b and the base's numerator r_ij r_ik + lambda d_a . d_b are
// formed in FP64 (from FP32 displacements the numerator would lose ~1e-6
// absolute, which zeta = 16 turns into 1e-5 relative error of a small row);
// everything after them -- exp, log, powers, cutoff, products -- is FP32.  exp(-eta (r_ij^2 + r_ik^2 + r_jk^2)) factors into the two
// neighbours' stored terms and one exp of r_jk.
__device__ __forceinline__ float uniform_f(double v)
{
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int((float)v)));
}

__device__ __forceinline__ double uniform_d(double v)
{
    const unsigned long long u = __double_as_longlong(v);
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)u);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(u >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// wave-uniform parameters (scalar registers) of the FP32-bulk path
struct Fast {
    float zeta[4];
    double lp[2], lm[2], lam[2];  // the base's numerator is FP64
    float eta[2];                 // eta_e
    float etal[2];                // eta_e log2(e)
    double rq0;  // rc^2 of species pair 0 (the pair enumeration's test for one cutoff)
};

// The zeta grid of the FP32-bulk kernels: ZG = 0 any grid (B^zeta =
// exp2(zeta log2 B)); ZG = 1 and 2 the power-of-two grids (1, 2, 4, 16) and
// (1, 2, 4, 8), in that order, whose powers come from a squaring chain with
// the exponents known at compile time (no transcendental per power).
template <int ZG>
struct ZGrid {
    static constexpr int e[4] = {0, 0, 0, 0};
    static constexpr int nsq = 0;
};
template <>
struct ZGrid<1> {
    static constexpr int e[4] = {0, 1, 2, 4};
    static constexpr int nsq = 4;
};
template <>
struct ZGrid<2> {
    static constexpr int e[4] = {0, 1, 2, 3};
    static constexpr int nsq = 3;
};

// B^(2^k), k = 0..nsq
template <int ZG>
__device__ __forceinline__ void squares(float b, float (&sq)[5])
{
    sq[0] = b;
#pragma unroll
    for (int k = 1; k < 5; ++k) sq[k] = k <= ZGrid<ZG>::nsq ? sq[k - 1] * sq[k - 1] : 0.0f;
}

// B^(2^k - 1), k = 0..nsq: 1, B, B^3, B^7, B^15
template <int ZG>
__device__ __forceinline__ void squares_m1(const float (&sq)[5], float (&mk)[5])
{
    mk[0] = 1.0f;
#pragma unroll
    for (int k = 1; k < 5; ++k) mk[k] = k <= ZGrid<ZG>::nsq ? mk[k - 1] * sq[k - 1] : 0.0f;
}

// The powers of the power-of-two zeta grids for the FP64 form: B^zeta_z
// (pw) or B^(zeta_z - 1) (pm) from the same squaring chain, multiplied in the
// same order as pow_bits (so bitwise its results), and 0 for B <= 0 as the
// reference's guard (src/SymmetryFunctions.cpp:121).
template <int ZG>
__device__ __forceinline__ void grid_powers64(double b, double (&pw)[4])
{
    const double bc = b > 0.0 ? b : 0.0;
    double sq[5];
    sq[0] = bc;
#pragma unroll
    for (int k = 1; k < 5; ++k) sq[k] = k <= ZGrid<ZG>::nsq ? sq[k - 1] * sq[k - 1] : 0.0;
#pragma unroll
    for (int z = 0; z < 4; ++z) pw[z] = sq[ZGrid<ZG>::e[z]];
}

template <int ZG>
__device__ __forceinline__ void grid_powers64_m1(double b, double (&pm)[4])
{
    const double bc = b > 0.0 ? b : 0.0;
    double sq[5], mk[5];
    sq[0] = bc;
#pragma unroll
    for (int k = 1; k < 5; ++k) sq[k] = k <= ZGrid<ZG>::nsq ? sq[k - 1] * sq[k - 1] : 0.0;
    mk[0] = 1.0;
#pragma unroll
    for (int k = 1; k < 5; ++k) mk[k] = k <= ZGrid<ZG>::nsq ? mk[k - 1] * sq[k - 1] : 0.0;
#pragma unroll
    for (int z = 0; z < 4; ++z) pm[z] = b > 0.0 ? mk[ZGrid<ZG>::e[z]] : 0.0;
}

// the species-pair index of two-species block p (0-0, 0-1, 1-1: 0, 1, 3)
__device__ __forceinline__ int blk_pair(int p) { return p + (p >> 1); }

template <int NPB>
__device__ __forceinline__ Fast load_fast(const Shared &S, const WaveParams &wp)
{
    Fast F;
#pragma unroll
    for (int z = 0; z < 4; ++z) F.zeta[z] = uniform_f(wp.zeta[z]);
#pragma unroll
    for (int l = 0; l < 2; ++l) {
        F.lp[l] = uniform_d(wp.lp[l]);
        F.lm[l] = uniform_d(wp.lm[l]);
        F.lam[l] = uniform_d(wp.lam[l]);
    }
#pragma unroll
    for (int e = 0; e < 2; ++e) {
        F.eta[e] = uniform_f(wp.eta3[e]);
        F.etal[e] = uniform_f(wp.eta3[e] * 1.4426950408889634);
    }
    F.rq0 = uniform_d(S.rc2[0]);
    return F;
}

// sin(y) and cos(y) on [0, pi/2] from their Taylor series to y^13 / y^12 in
// Horner form (truncation below 7e-9)
__device__ __forceinline__ float sin_q(float y, float y2)
{
    float t = fmaf(y2, 1.0f / 6227020800.0f, -1.0f / 39916800.0f);
    t = fmaf(y2, t, 1.0f / 362880.0f);
    t = fmaf(y2, t, -1.0f / 5040.0f);
    t = fmaf(y2, t, 1.0f / 120.0f);
    t = fmaf(y2, t, -1.0f / 6.0f);
    return fmaf(y * y2, t, y);
}

__device__ __forceinline__ float cos_q(float y2)
{
    float t = fmaf(y2, 1.0f / 479001600.0f, -1.0f / 3628800.0f);
    t = fmaf(y2, t, 1.0f / 40320.0f);
    t = fmaf(y2, t, -1.0f / 720.0f);
    t = fmaf(y2, t, 1.0f / 24.0f);
    t = fmaf(y2, t, -0.5f);
    return fmaf(y2, t, 1.0f);
}

// |e|^2 of e = d_b - d_a; the enumeration and the triplet phase evaluate the
// same expression, so an active pair has r_jk^2 < rc^2 in both
__device__ __forceinline__ double rjk2_of(const double2 &a0, const double2 &a1, const double2 &b0,
                                          const double2 &b1, double &ex, double &ey, double &ez)
{
    ex = b0.x - a0.x;
    ey = b0.y - a0.y;
    ez = b1.x - a1.x;
    return fma(ez, ez, fma(ey, ey, ex * ex));
}

// geometry of an active pair (a, b) of the FP32-bulk path
struct TripF {
    double ex, ey, ez;  // d_b - d_a
    double rjk2d;
    float rjk2, rjk, rjkinv;
    float sn, y;        // sin(y), y = (pi/2)(rc - r_jk)/rc: fc_jk = sin^2 y
};

__device__ __forceinline__ void geom_f(const double2 &a0, const double2 &a1, const double2 &b0,
                                       const double2 &b1, float rc, double rc2, float rcs,
                                       TripF &t)
{
    t.rjk2d = rjk2_of(a0, a1, b0, b1, t.ex, t.ey, t.ez);
    t.rjk2 = (float)t.rjk2d;
    // (coincident neighbours, r_jk = 0: a finite 1 / r_jk and r_jk = 0, as the
    // reference's forward, whose fc_jk(0) = 1)
    t.rjkinv = __builtin_amdgcn_rsqf(fmaxf(t.rjk2, 1e-30f));
    t.rjk = t.rjk2 * t.rjkinv;
    // rc - r_jk = (rc^2 - r_jk^2) / (rc + r_jk) keeps its relative accuracy near rc
    t.y = (float)(rc2 - t.rjk2d) * __builtin_amdgcn_rcpf(rc + t.rjk) * rcs;
    t.sn = sin_q(t.y, t.y * t.y);
}

// the base 1 + lambda_l cos(theta) = (r_ij r_ik + lambda_l d_j . d_k) /
// (r_ij r_ik), numerator in FP64 from the FP64 displacements (rr = r_ij r_ik,
// dot = d_j . d_k; for lambda = -1 and a narrow angle the two nearly cancel,
// which FP64 carries to ~1e-15 of r_ij r_ik), inv1 = 1 / (r_ij r_ik) in FP32;
// a base <= 0 becomes 0, and its powers and their derivatives 0 as the
// reference's base <= 0 branch (src/SymmetryFunctions.cpp:118-119)
__device__ __forceinline__ float base_f(const Fast &F, int l, double rr, double dot, float inv1)
{
    return fmaxf((float)fma(F.lam[l], dot, rr) * inv1, 0.0f);
}

// the same base as (lp (r_ij + r_ik)^2 - lm (r_ij - r_ik)^2 - lambda r_jk^2) /
// (2 r_ij r_ik) (sp, sm: the two squares; inv = 1 / (2 r_ij r_ik)): the
// forward kernel's form, which holds fewer registers
__device__ __forceinline__ float base_sq_f(const Fast &F, int l, double sp, double sm, double rjk2,
                                           float inv)
{
    const double num = fma(-F.lam[l], rjk2, fma(F.lp[l], sp, -F.lm[l] * sm));
    return fmaxf((float)num * inv, 0.0f);
}

// d_a . d_b of two compacted neighbours of the FP32-bulk path (D layout)
__device__ __forceinline__ double dot_ab(const double2 &a0, const double2 &a1, const double2 &b0,
                                         const double2 &b1)
{
    return fma(a1.x, b1.x, fma(a0.y, b0.y, a0.x * b0.x));
}

// The g2 rows of one neighbour at distance r in the FP32-bulk form, with
// phi_q = exp(-eta_q (r - Rs_q)^2) fc(r).  ADJ: returns sum_q dE/dzeta_q
// d(phi_q)/dr (weights gw[8 sj + q], FP32, in LDS).  FWD: writes the terms
// phi_q to term[0..8) (16-byte aligned LDS).  The exponent is formed in FP32:
// relative error about 3e-7 |eta (r - Rs)^2| per term, the largest terms
// having the smallest exponents.  Rows past the set's g2 rows have eta 0,
// weight 0 and no slot.
template <int NSB, bool ADJ, bool FWD>
__device__ __forceinline__ float pair_rows_f(const WaveParams &wp, const float *gw, float r,
                                             float fc, float dfc, int sj, float *term)
{
    float gq[8];
    if (ADJ) {
        const float2 *G = reinterpret_cast<const float2 *>(gw + (NSB == 1 ? 0 : 8 * sj));
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const float2 v = G[k];
            gq[2 * k] = v.x;
            gq[2 * k + 1] = v.y;
        }
    }
    float sum = 0.0f, t[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const float dr = r - wp.rs2f[q];
        const float ex = __builtin_amdgcn_exp2f(-wp.e2l[q] * (dr * dr));
        if (ADJ) sum = fmaf(gq[q], ex * fmaf(wp.m2e[q] * dr, fc, dfc), sum);
        t[q] = ex * fc;
    }
    if (FWD) {
        float4 *T4 = reinterpret_cast<float4 *>(term);
        T4[0] = make_float4(t[0], t[1], t[2], t[3]);
        T4[1] = make_float4(t[4], t[5], t[6], t[7]);
    }
    return sum;
}

// Phase 2: the active pairs (a < b, r_jk < rc_jk) of the M compacted
// neighbours, walked in rotational order, compacted into one region of
// L.l0 per species-pair block (codes a | b << 6).  With two species and n1
// neighbours of species 1, block 0 (0-0), 1 (0-1) and 2 (1-1) hold at most
// n0(n0-1)/2, n0 n1 and n1(n1-1)/2 pairs, so the regions start at 0,
// n0(n0-1)/2 and n0(n0-1)/2 + n0 n1.  Two species blocks come with the slots
// sorted by species (species 0 at 0 .. n0 - 1), so a pair's block follows
// from its two indices.  UNI: every species pair has the cutoff of pair 0
// (no per-pair choice).  The neighbours sit at slots off .. off + M - 1
// (off = 0, or 32 for the second centre of a paired wave); the stored codes
// carry the slots relative to off.
template <int P, int NPB, bool UNI>
__device__ __forceinline__ void enumerate_pairs(const Shared &S, const Grp &L, const Fast &F,
                                                int ns, int M, int n1, int s, int (&cnt)[3],
                                                int (&base)[3], int off = 0)
{
    const int n0 = M - n1;
    base[0] = 0;
    base[1] = (n0 * (n0 - 1)) >> 1;
    base[2] = base[1] + n0 * n1;
    cnt[0] = cnt[1] = cnt[2] = 0;
    const int T = (M * (M - 1)) >> 1;
    const float inv_m = __builtin_amdgcn_rcpf((float)(M > 1 ? M : 2));
    for (int t0 = 0; t0 < T; t0 += kW) {
        const int t = t0 + s;
        // no branch on t < T: past T the lane takes pair 0, whose reads stay
        // inside the arrays, and the test is masked after, so that the
        // ballots below take the compare results as they are
        int a, b;
        pair_of<NPB == 1>(t < T ? t : 0, M, inv_m, a, b);
        const int code = a | (b << 6);
        // species of the two slots: from the indices with two blocks
        const bool xa = a >= n0, xb = b >= n0;
        int sa, sb;
        bool act;
        if constexpr (P == 0) {
            const double *px = L.px + off, *py = L.py + off, *pz = L.pz + off;
            const double rjk2 = dist2(px[b] - px[a], py[b] - py[a], pz[b] - pz[a]);
            sa = NPB > 1 ? (int)xa : L.sp[off + a];
            sb = NPB > 1 ? (int)xb : L.sp[off + b];
            // sqrt(rjk2) < rc_jk as the reference tests it, without the root
            act = rjk2 < S.rct[UNI ? 0 : pidx(sa, sb, ns)];
        } else {
            // the triplet phase forms the same r_jk^2 from the same operands:
            // an active pair has r_jk^2 < rc^2 there too.  (The reference
            // tests r_jk < rc_jk; the two differ only within an ulp of rc_jk,
            // where fc_jk vanishes quadratically.)
            const double2 *D0 = reinterpret_cast<const double2 *>(L.px);
            const float4 *V = reinterpret_cast<const float4 *>(D0 + 2 * kW) + off;
            const double2 *D = D0 + off;
            double ex, ey, ez;
            const double rjk2 = rjk2_of(D[a], D[kW + a], D[b], D[kW + b], ex, ey, ez);
            sa = NPB > 1 ? (int)xa : __float_as_int(V[a].w);
            sb = NPB > 1 ? (int)xb : __float_as_int(V[b].w);
            const double rc2 = (UNI || ns == 1) ? F.rq0
                               : (NPB > 1 ? S.rc2[blk_pair(sa + sb)] : S.rc2[pidx(sa, sb, ns)]);
            act = rjk2 < rc2;
        }
        // the lanes past the end of the list, as a scalar mask (a ballot of
        // act && t < T would be rebuilt from a per-lane boolean)
        const unsigned long long live = T - t0 >= kW ? ~0ull : (1ull << (T - t0)) - 1ull;
        const unsigned long long mt = __builtin_amdgcn_ballot_w64(act) & live;
        act = act && t < T;
        if (NPB == 1) {
            const unsigned long long m = mt;
            if (act) L.l0[cnt[0] + below(m, s)] = (unsigned short)code;
            cnt[0] += __popcll(m);
        } else {
            // two species: blocks 0-0, 0-1, 1-1 = pair_index(2, sa, sb); the
            // block masks as scalar logic on the two index compares' ballots (a
            // ballot of act && blk == p would be rebuilt from a per-lane boolean)
            const unsigned long long ma = mt;
            const unsigned long long ba = __builtin_amdgcn_ballot_w64(xa);
            const unsigned long long bb = __builtin_amdgcn_ballot_w64(xb);
            const unsigned long long m0 = ma & ~(ba | bb), m1 = ma & (ba ^ bb), m2 = ma & ba & bb;
            if (act) {
                const int p0 = base[0] + cnt[0] + below(m0, s);
                const int p1 = base[1] + cnt[1] + below(m1, s);
                const int p2 = base[2] + cnt[2] + below(m2, s);
                L.l0[xa != xb ? p1 : (xa ? p2 : p0)] = (unsigned short)code;
            }
            cnt[0] += __popcll(m0);
            cnt[1] += __popcll(m1);
            cnt[2] += __popcll(m2);
        }
    }
}

// shared quantities of one active triplet (a, b) of the FP64 form
template <int P>
struct Trip {
    using T = typename Real<P>::T;
    double ra, rb, ria, rib, rjk, rjkinv;
    double e0, e1, e2;  // x_k - x_j
    T fa, fb, fjk, dfjk;
    T B[2], E[2];
    T inv2;           // 1 / (2 r_ij r_ik)
    double nij, nik;  // r_ij^2 - r_ik^2 + r_jk^2 and r_ik^2 - r_ij^2 + r_jk^2
};

// sym_g4's own arithmetic and operation order in FP64
// (src/SymmetryFunctions.cpp:101-127).  The FP64 form keeps it, the cutoff
// 0.5 (cos(pi r_jk / rc) + 1) included: near rc that value carries ~1e-16
// absolute rounding, and a species block whose value one such triplet
// dominates must reproduce it to stay within 1e-12 of the reference.
template <int P, bool ADJ>
__device__ __forceinline__ void triplet(const Shared &S, const WaveParams &wp, const WaveTab &wt,
                                        const Grp &L, int ns, int a, int b, Trip<P> &t)
{
    static_assert(P == 0, "the FP32-bulk form takes adj_triplet_f and the forward's own path");
    using T = typename Real<P>::T;
    const int k = pidx(L.sp[a], L.sp[b], ns);
    const double rc = S.rc[k], rcinv = S.rcinv[k];
    t.ra = L.r[a];
    t.rb = L.r[b];
    t.ria = L.rinv[a];
    t.rib = L.rinv[b];
    t.fa = (T)L.fc[a];
    t.fb = (T)L.fc[b];
    // x_k - x_j from the absolute coordinates, as src/SymmetryFunctions.cpp:216-217
    t.e0 = L.px[b] - L.px[a];
    t.e1 = L.py[b] - L.py[a];
    t.e2 = L.pz[b] - L.pz[a];
    const double rjk2 = dist2(t.e0, t.e1, t.e2);
    t.B[0] = t.B[1] = (T)0;
    t.E[0] = t.E[1] = (T)0;
    t.rjk = sqrt(rjk2);
    const RefTrip rt = ref_trip(t.ra, t.rb, t.rjk);
#pragma unroll
    for (int l = 0; l < 2; ++l)
        if (l < wt.nl) t.B[l] = (T)ref_base(wp.lam[l], rt.cosv);
#pragma unroll
    for (int e = 0; e < 2; ++e)
        if (e < wt.ne) t.E[e] = (T)exp(-wp.eta3[e] * rt.s2);
    double sn, cs;
    sincos(ref_arg(t.rjk, rc), &sn, &cs);
    t.fjk = (T)(0.5 * (cs + 1.0));
    t.dfjk = (T)(-0.5 * sn * (kPi * rcinv));
    const double h = 0.5 * t.ria * t.rib;
    t.inv2 = (T)h;
    if (ADJ) {
        t.rjkinv = 1.0 / t.rjk;
        t.nij = rt.rijsq - rt.riksq + rt.rjksq;
        t.nik = rt.riksq - rt.rijsq + rt.rjksq;
    }
}

// One active triplet of the FP32-bulk adjoint: the derivatives of the row
// sums with respect to r_ij, r_ik and r_jk, turned into vectors and added
// into the two neighbours' FP64 LDS accumulators.  w / wv: the block's row
// weights dE/dzeta 2^(1-zeta) and those times zeta lambda.  FWD: the triplet's
// descriptor terms B^zeta exp(..) fc fc fc (without 2^(1-zeta)) are also added
// into the lane's row accumulators `acc`, from the same powers.  a and b are
// slots relative to off (the centre's first slot).
template <int NPB, bool FWD, int ZG>
__device__ __forceinline__ void adj_triplet_f(const Shared &S, const Fast &F, const Grp &L, int ns,
                                              float rc, double rc2, float rcs, int off, int a, int b,
                                              const float *w, const float *wv, double (&acc)[16])
{
    const double2 *D0 = reinterpret_cast<const double2 *>(L.px);
    const float4 *V0 = reinterpret_cast<const float4 *>(D0 + 2 * kW);
    const float2 *W = reinterpret_cast<const float2 *>(V0 + kW) + off;
    const double2 *D = D0 + off;
    const float4 *V = V0 + off;
    double *fx = L.fx + off, *fy = L.fy + off, *fz = L.fz + off;
    const double2 a0 = D[a], a1 = D[kW + a], b0 = D[b], b1 = D[kW + b];
    const float4 va = V[a], vb = V[b];
    const float2 ea = W[a], eb = W[b];
    if (NPB == 1 && ns > 1) {  // one block, cutoff per species pair
        const int k = pidx(__float_as_int(va.w), __float_as_int(vb.w), ns);
        rc = (float)S.rc[k];
        rc2 = S.rc2[k];
        rcs = (float)S.rcs[k];
    }
    TripF g;
    geom_f(a0, a1, b0, b1, rc, rc2, rcs, g);
    const float cs = cos_q(g.y * g.y);
    const float fjk = g.sn * g.sn, dfjk = -2.0f * rcs * g.sn * cs;
    // r in FP32: beside 1/r in the adjoint kernel's two-block lists; the fused
    // kernel, short of registers, converts it at its use
    const bool vr = NPB > 1 && !FWD;
    const float ra = vr ? va.w : (float)a1.y, rb = vr ? vb.w : (float)b1.y;
    const float ria = va.x, rib = vb.x;
    const float fa = va.y, fb = vb.y, dfa = va.z, dfb = vb.z;
    const float inv1 = ria * rib;  // 1 / (r_ij r_ik)
    const double rr = a1.y * b1.y, dot = dot_ab(a0, a1, b0, b1);
    // the block's row weights: the adjoint reads all 32 up front, four to an
    // LDS read; the fused kernel, short of registers, two at their use
    float wl[16], wvl[16];
    if constexpr (!FWD) {
        const float4 *w4 = reinterpret_cast<const float4 *>(__builtin_assume_aligned(w, 16));
        const float4 *v4 = reinterpret_cast<const float4 *>(__builtin_assume_aligned(wv, 16));
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const float4 a = w4[k], b = v4[k];
            wl[4 * k] = a.x;
            wl[4 * k + 1] = a.y;
            wl[4 * k + 2] = a.z;
            wl[4 * k + 3] = a.w;
            wvl[4 * k] = b.x;
            wvl[4 * k + 1] = b.y;
            wvl[4 * k + 2] = b.z;
            wvl[4 * k + 3] = b.w;
        }
    }
    float E[2];
    E[0] = ea.x * eb.x * __builtin_amdgcn_exp2f(-F.etal[0] * g.rjk2);
    E[1] = ea.y * eb.y * __builtin_amdgcn_exp2f(-F.etal[1] * g.rjk2);
    float U[2] = {0.0f, 0.0f}, Vs[2] = {0.0f, 0.0f};
    double te[2] = {0.0, 0.0};  // FWD: FP32 terms, summed in FP64
    if (FWD) {
        const float fc3f = fa * fb * fjk;
        te[0] = (double)(E[0] * fc3f);
        te[1] = (double)(E[1] * fc3f);
    }
#pragma unroll
    for (int l = 0; l < 2; ++l) {
        const float bl = base_f(F, l, rr, dot, inv1);
        // power-of-two grids: B^(2^k) and B^(2^k - 1) by squaring; else
        // B^(zeta - 1) = exp2((zeta - 1) log2 B)
        // (B = 0: B^0 = 0 as well, the derivative of the reference's 0 branch)
        float sq[5], mk[5], lb = 0.0f;
        if constexpr (ZG != 0) {
            squares<ZG>(bl, sq);
            squares_m1<ZG>(sq, mk);
            mk[0] = bl > 0.0f ? 1.0f : 0.0f;
        } else {
            lb = __builtin_amdgcn_logf(fmaxf(bl, 1e-30f));
        }
#pragma unroll
        for (int z = 0; z < 4; ++z) {
            const float pm = ZG != 0 ? mk[ZGrid<ZG>::e[z]]  // B^(zeta - 1)
                                     : (bl > 0.0f ? __builtin_amdgcn_exp2f((F.zeta[z] - 1.0f) * lb)
                                                  : 0.0f);
            const float pw = ZG != 0 ? sq[ZGrid<ZG>::e[z]] : pm * bl;
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                const int k = (z * 2 + l) * 2 + e;
                U[e] = fmaf(FWD ? w[k] : wl[k], pw, U[e]);
                Vs[e] = fmaf(FWD ? wv[k] : wvl[k], pm, Vs[e]);
                if (FWD) acc[(z * 2 + l) * 2 + e] = fma((double)pw, te[e], acc[(z * 2 + l) * 2 + e]);
            }
        }
    }
    const float ue0 = U[0] * E[0], ue1 = U[1] * E[1];
    const float SA = fmaf(Vs[0], E[0], Vs[1] * E[1]);
    const float SB = ue0 + ue1;
    const float SC = fmaf(F.eta[0], ue0, F.eta[1] * ue1);
    // d cos(theta) / d r_ij = (r_ij^2 - r_ik^2 + r_jk^2) / (2 r_ij^2 r_ik), whose
    // numerator is 2 (r_ij^2 - d_a . d_b); likewise for r_ik with 2 (r_ik^2 - d_a . d_b)
    const float nij = (float)fma(a1.y, a1.y, -dot), nik = (float)fma(b1.y, b1.y, -dot);
    const float dcij = nij * inv1 * ria, dcik = nik * inv1 * rib, dcjk = -g.rjk * inv1;
    const float fc3 = fa * fb * fjk;
    const float Aij = fmaf(fc3, fmaf(dcij, SA, -2.0f * ra * SC), fb * fjk * dfa * SB);
    const float Aik = fmaf(fc3, fmaf(dcik, SA, -2.0f * rb * SC), fa * fjk * dfb * SB);
    const float Ajk = fmaf(fc3, fmaf(dcjk, SA, -2.0f * g.rjk * SC), fa * fb * dfjk * SB);
    const double a1c = (double)(Aij * ria), a2c = (double)(Aik * rib), a3c = (double)(Ajk * g.rjkinv);
    // the r_jk term once, added to b and (negated, exactly) to a
    const double jx = a3c * g.ex, jy = a3c * g.ey, jz = a3c * g.ez;
    atomicAdd(&fx[a], fma(a1c, a0.x, -jx));
    atomicAdd(&fy[a], fma(a1c, a0.y, -jy));
    atomicAdd(&fz[a], fma(a1c, a1.x, -jz));
    atomicAdd(&fx[b], fma(a2c, b0.x, jx));
    atomicAdd(&fy[b], fma(a2c, b0.y, jy));
    atomicAdd(&fz[b], fma(a2c, b1.x, jz));
}

// ZG: the zeta grid of the FP32-bulk form (ZGrid)
// DC: the launch uses DevCounts (the serving loop's device-side sizes,
// per-request flags and per-centre rows); without, every DevCounts branch
// and pointer is compiled out.
template <int P, int NSB, int NPB, int ZG, bool DC>
__global__ __launch_bounds__(64, P == 0 ? 3 : 5) void k_fwd_wave(WaveTab wt, int n_centers,
                                                 const int *__restrict__ centers,
                                                 const int *__restrict__ offs,
                                                 const int *__restrict__ nl,
                                                 const int *__restrict__ species,
                                                 const double *__restrict__ x,
                                                 double *__restrict__ desc, int width, int maxn,
                                                 DevCounts dc_in)
{
    using T = typename Real<P>::T;
    const Real<P> R;
    const DevCounts dc = DC ? dc_in : DevCounts();
    extern __shared__ double lds_base[];
    const int s = threadIdx.x;
    const int ns = NSB == 2 ? 2 : wt.n_species;  // resolved: two species
    const Shared S = carve_shared(lds_base, ns);
    const Grp L = carve_group(lds_base, ns, kFwd);
    stage_shared(S, wt, s);
    const WaveParams &wp = *S.par;
    const int nsq = __builtin_amdgcn_readfirstlane(wp.nsq);
    const Fast F = P == 1 ? load_fast<NPB>(S, wp) : Fast();
    const bool assign = dc.out_row != nullptr;
    if (dc.hdr) n_centers = dc.hdr[0];
    const TileRange tr = tile_range(n_centers);
    if (tr.first >= tr.hi) return;
    // entry s of a centre's descriptor row before this call: += keeps it
    // (the reference's semantics), the serving loop's = starts from zero
    auto row_in = [&](const Ctr &C) -> double {
        return (!assign && C.work && s < width) ? desc[(size_t)C.c * width + s] : 0.0;
    };
    // the first centre's loads, in order
    Ctr C = load_centre(tr.first, n_centers, centers, offs, species, x, dc);
    double rin = row_in(C);
    Nbr B = load_nbr(load_index(C, s, nl), species, x, wt.nl_species != 0);
    for (int c = tr.first; c < tr.hi; c += tr.step) {
        const Ctr Cn = load_centre(c + tr.step, n_centers, centers, offs, species, x, dc);
        const double rinn = row_in(Cn);
        double *row = C.work ? (dc.out_row ? dc.out_row[c] : desc + (size_t)c * width) : nullptr;
        // the row is summed in LDS and stored once, coalesced, at the end
        if (s < kWaveSlots) L.d[s] = rin;
        __syncthreads();

        // ---- pairs ----
        bool valid = false;
        double d0 = 0, d1 = 0, d2 = 0, r2 = 0, r = 0;
        const int sj = B.sj;
        if (s < C.N) {
            d0 = B.p0 - C.x0;
            d1 = B.p1 - C.x1;
            d2 = B.p2 - C.x2;
            r2 = dist2(d0, d1, d2);
            r = sqrt(r2);
            valid = r < S.rc[pidx(C.si, sj, ns)];  // r > rc is skipped; r == rc adds exactly 0
        }
        const unsigned long long m = __ballot(valid);
        const int M = __popcll(m);
        // two species blocks: species-0 neighbours first, list order within each
        const unsigned long long m1 = NSB == 2 ? __ballot(valid && sj == 1) : 0ull;
        const int n1 = NSB == 2 ? __popcll(m1) : 0;
        T *terms = reinterpret_cast<T *>(L.l0);  // the g2 terms, until the enumeration
        if (valid) {
            const int pos = NSB == 2 ? (sj ? (M - n1) + below(m1, s) : below(m & ~m1, s))
                                     : below(m, s);
            const int k = pidx(C.si, sj, ns);
            T fc, dfc;
            cut_pair(r, S.rc[k], S.rcinv[k], S.rcs[k], fc, dfc, R);
            if constexpr (P == 1) {
                double2 *D = reinterpret_cast<double2 *>(L.px);
                float4 *V = reinterpret_cast<float4 *>(D + 2 * kW);
                D[pos] = make_double2(d0, d1);
                D[kW + pos] = make_double2(d2, r);
                const float r2f = (float)r2;
                V[pos] = make_float4(0.70710678f * rcp_refined((float)r),
                                     (float)fc * __builtin_amdgcn_exp2f(-F.etal[0] * r2f),
                                     (float)fc * __builtin_amdgcn_exp2f(-F.etal[1] * r2f),
                                     __int_as_float(sj));
                pair_rows_f<NSB, false, true>(wp, nullptr, (float)r, fc, 0.0f, sj,
                                              reinterpret_cast<float *>(terms) + 8 * pos);
            } else {
                L.px[pos] = B.p0;
                L.py[pos] = B.p1;
                L.pz[pos] = B.p2;
                L.r[pos] = r;
                L.rinv[pos] = 1.0 / r;
                L.fc[pos] = (double)fc;
                L.sp[pos] = sj;
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    double v = 0.0;
                    if (q < wt.nr2) {
                        const double dr = r - wp.rs2[q];
                        v = (double)(dexp(-wp.eta2[q] * dr * dr, R) * fc);
                    }
                    terms[8 * pos + q] = (T)v;
                }
            }
        }
        const int jn = load_index(Cn, s, nl);
        __syncthreads();
        g2_rows_sum<NSB>(terms, M, M - n1, s, wp.slots, L.d);
        __syncthreads();

        // ---- triplets ----
        Nbr Bn;
        if (wt.has3) {
            int cnt[3], base[3];
            if (wt.rc_uniform)
                enumerate_pairs<P, NPB, true>(S, L, F, ns, M, n1, s, cnt, base);
            else
                enumerate_pairs<P, NPB, false>(S, L, F, ns, M, n1, s, cnt, base);
            Bn = load_nbr(jn, species, x, wt.nl_species != 0);
            __syncthreads();
            // the blocks are dealt columns of 4 lanes (s, s ^ 16, s ^ 32, s ^ 48;
            // column s & 15) in proportion to their active pairs: the fewest
            // passes I with sum_p ceil(cnt_p / 4I) <= 16 columns.  Lane s serves
            // block p with stride 4 * cols(p); a column's 4 lanes sum with two
            // register swaps, the columns of a block in LDS.
            const int nall = cnt[0] + (NPB > 1 ? cnt[1] + cnt[2] : 0);
            int passes = (nall + kW - 1) / kW, nc[3] = {16, 0, 0};
            if (NPB > 1 && nall > 0) {
                // ceil(c / 4I) = ceil(ceil(c / 4) / I)
                const int c40 = (cnt[0] + 3) >> 2, c41 = (cnt[1] + 3) >> 2, c42 = (cnt[2] + 3) >> 2;
                auto cols_for = [&](int I) {
                    nc[0] = ceil_div_small(c40, I);
                    nc[1] = ceil_div_small(c41, I);
                    nc[2] = ceil_div_small(c42, I);
                    return nc[0] + nc[1] + nc[2];
                };
                while (cols_for(passes) > 16) ++passes;
            }
            const int col = s & 15, rowq = s >> 4;
            const int pl = NPB == 1 ? 0 : (col < nc[0] ? 0 : (col < nc[0] + nc[1] ? 1 : 2));
            const int cs = pl == 0 ? 0 : (pl == 1 ? nc[0] : nc[0] + nc[1]);
            const int ncol = pl == 0 ? nc[0] : (pl == 1 ? nc[1] : nc[2]);
            const int u = NPB == 1 ? s : (col - cs) + rowq * ncol;
            const int stride = NPB == 1 ? kW : 4 * ncol;
            const int nb = NPB == 1 ? cnt[0] : (pl == 0 ? cnt[0] : (pl == 1 ? cnt[1] : cnt[2]));
            const unsigned short *lst =
                L.l0 + (NPB == 1 ? 0 : (pl == 0 ? base[0] : (pl == 1 ? base[1] : base[2])));
            const bool mine = NPB == 1 || col < cs + ncol;
            {
                double v4[4];
                if constexpr (P == 1) {
                    // this lane's species-pair cutoff
                    // the lane's block cutoff, from LDS (one read per centre)
                    const int kp = blk_pair(pl);
                    const float rcl = (float)S.rc[kp];
                    const double rc2l = S.rc2[kp];
                    const float rcsl = (float)S.rcs[kp];
                    const bool rc_pair = NPB == 1 && ns > 1;  // cutoff per species pair
                    const double2 *D = reinterpret_cast<const double2 *>(L.px);
                    const float4 *V = reinterpret_cast<const float4 *>(D + 2 * kW);
                    // FP32 terms, FP64 sums (per lane, then over lanes and columns)
                    double acc[16];
#pragma unroll
                    for (int k = 0; k < 16; ++k) acc[k] = 0.0;
                    for (int i = 0; i < passes; ++i) {
                        const int t = u + i * stride;
                        if (mine && t < nb) {
                            const int code = lst[t];
                            const int a = code & 63, b = (code >> 6) & 63;
                            const double2 a0 = D[a], a1 = D[kW + a], b0 = D[b], b1 = D[kW + b];
                            const float4 va = V[a], vb = V[b];
                            float rc = rcl, rcs = rcsl;
                            double rc2 = rc2l;
                            if (rc_pair) {
                                const int k = pidx(__float_as_int(va.w), __float_as_int(vb.w), ns);
                                rc = (float)S.rc[k];
                                rc2 = S.rc2[k];
                                rcs = (float)S.rcs[k];
                            }
                            TripF g;
                            geom_f(a0, a1, b0, b1, rc, rc2, rcs, g);
                            const float fjk = g.sn * g.sn;
                            const float inv = va.x * vb.x;  // 1 / (2 r_ij r_ik)
                            const double sp = (a1.y + b1.y) * (a1.y + b1.y);
                            const double sm = (a1.y - b1.y) * (a1.y - b1.y);
                            double te[2];
                            te[0] = (double)(va.y * vb.y * (fjk * __builtin_amdgcn_exp2f(-F.etal[0] * g.rjk2)));
                            te[1] = (double)(va.z * vb.z * (fjk * __builtin_amdgcn_exp2f(-F.etal[1] * g.rjk2)));
#pragma unroll
                            for (int l = 0; l < 2; ++l) {
                                const float bl = base_sq_f(F, l, sp, sm, g.rjk2d, inv);
                                float sq[5], lb = 0.0f;
                                if constexpr (ZG != 0) squares<ZG>(bl, sq);
                                else lb = __builtin_amdgcn_logf(fmaxf(bl, 1e-30f));
#pragma unroll
                                for (int z = 0; z < 4; ++z) {
                                    const double pw =
                                        ZG != 0 ? (double)sq[ZGrid<ZG>::e[z]]
                                                : (bl > 0.0f ? (double)__builtin_amdgcn_exp2f(F.zeta[z] * lb)
                                                             : 0.0);
#pragma unroll
                                    for (int e = 0; e < 2; ++e)
                                        acc[(z * 2 + l) * 2 + e] = fma(pw, te[e], acc[(z * 2 + l) * 2 + e]);
                                }
                            }
                        }
                    }
                    // a column's 4 lanes (value halves swapped between lane
                    // halves, then rows), in FP64: lane s keeps values 4 (s >> 4) .. +3
#pragma unroll
                    for (int q = 0; q < 8; ++q) {
                        double a = acc[q], b = acc[q + 8];
                        xswap<32>(a, b);
                        acc[q] = a + b;
                    }
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        double a = acc[q], b = acc[q + 4];
                        xswap<16>(a, b);
                        acc[q] = a + b;
                    }
#pragma unroll
                    for (int q = 0; q < 4; ++q) v4[q] = acc[q];
                } else {
                    double acc3[16];
#pragma unroll
                    for (int k = 0; k < 16; ++k) acc3[k] = 0.0;
                    for (int i = 0; i < passes; ++i) {
                        const int t = u + i * stride;
                        if (mine && t < nb) {
                            const int code = lst[t];
                            Trip<P> q;
                            triplet<P, false>(S, wp, wt, L, ns, code & 63, (code >> 6) & 63, q);
                            const T fc3 = q.fa * q.fb * q.fjk;
                            double te[2];
#pragma unroll
                            for (int e = 0; e < 2; ++e) te[e] = (double)(q.E[e] * fc3);
                            // the grid's own size: l < nl, z < nz, e < ne
#pragma unroll
                            for (int l = 0; l < 2; ++l) {
                                if (l < wt.nl) {
                                    T pw[4];
                                    if constexpr (P == 0 && ZG != 0)
                                        grid_powers64<ZG>(q.B[l], pw);
                                    else
                                        powers<P>(q.B[l], nsq, wp, wt.nz, pw);
#pragma unroll
                                    for (int z = 0; z < 4; ++z) {
                                        if (z < wt.nz) {
                                            const double pz = (double)((T)wp.coef[z] * pw[z]);
#pragma unroll
                                            for (int e = 0; e < 2; ++e)
                                                if (e < wt.ne)
                                                    acc3[(z * 2 + l) * 2 + e] =
                                                        fma(pz, te[e], acc3[(z * 2 + l) * 2 + e]);
                                        }
                                    }
                                }
                            }
                        }
                    }
                    // the column's 4 lanes: lane s keeps values 4 (s >> 4) .. +3
#pragma unroll
                    for (int q = 0; q < 8; ++q) {
                        double a = acc3[q], b = acc3[q + 8];
                        xswap<32>(a, b);
                        acc3[q] = a + b;
                    }
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        double a = acc3[q], b = acc3[q + 4];
                        xswap<16>(a, b);
                        acc3[q] = a + b;
                    }
#pragma unroll
                    for (int q = 0; q < 4; ++q) v4[q] = acc3[q];
                }
                // the neighbour arrays are dead here: value k of column c at
                // buf[kColStr k + c], then lane (p, k) = (s >> 4, s & 15) adds its
                // block's columns in order
                double *buf = L.px;
#pragma unroll
                for (int q = 0; q < 4; ++q) buf[kColStr * (4 * rowq + q) + col] = v4[q];
                __syncthreads();
                if (rowq < NPB) {
                    const int c0 = rowq == 0 ? 0 : (rowq == 1 ? nc[0] : nc[0] + nc[1]);
                    const int c1 = c0 + (rowq == 0 ? nc[0] : (rowq == 1 ? nc[1] : nc[2]));
                    double sum = 0.0;
                    for (int c = c0; c < c1; ++c) sum += buf[kColStr * col + c];
                    const int slot = wp.slots[16 + s];
                    // the FP32-bulk rows take their 2^(1 - zeta) here
                    if (P == 1) sum *= wp.coef[col >> 2];
                    if (c1 > c0 && slot >= 0) L.d[slot] += sum;
                }
            }
        } else {
            Bn = load_nbr(jn, species, x, wt.nl_species != 0);
        }
        __syncthreads();
        if (row && s < width) row[s] = L.d[s];
        C = Cn;
        B = Bn;
        rin = rinn;
    }
}

// The adjoint kernel (the fused forward + adjoint is k_fused_pair).
// Occupancy floors (waves per SIMD): FP64 none (168 VGPRs: 3), FP32-bulk 4 (128 VGPRs).
template <int P, int NSB, int NPB, int ZG, bool DC>
__global__ __launch_bounds__(64, P == 0 ? 1 : 4)
void k_adj_wave(WaveTab wt, int n_centers, const int *__restrict__ centers,
                const int *__restrict__ offs, const int *__restrict__ nl,
                const int *__restrict__ species, const double *__restrict__ x,
                const double *__restrict__ dE, const int *__restrict__ inv_pos,
                double *__restrict__ edge, int n_edges, int width, int maxn, DevCounts dc_in)
{
    using T = typename Real<P>::T;
    const Real<P> R;
    const DevCounts dc = DC ? dc_in : DevCounts();
    extern __shared__ double lds_base[];
    const int s = threadIdx.x;
    const int ns = NSB == 2 ? 2 : wt.n_species;  // resolved: two species
    const Shared S = carve_shared(lds_base, ns);
    const Grp L = carve_group(lds_base, ns, kAdj);
    stage_shared(S, wt, s);
    const WaveParams &wp = *S.par;
    const int nsq = __builtin_amdgcn_readfirstlane(wp.nsq);
    const Fast F = P == 1 ? load_fast<NPB>(S, wp) : Fast();
    if (dc.hdr) {
        n_centers = dc.hdr[0];
        n_edges = dc.hdr[2];
    }
    const TileRange tr = tile_range(n_centers);
    if (tr.first >= tr.hi) return;
    // dE/dzeta entry s of a centre's row, and the record position of the centre
    auto grad_of = [&](const Ctr &C) -> double {
        if (!C.work || s >= width) return 0.0;
        return dc.in_row ? dc.in_row[C.c][s] : dE[(size_t)C.c * width + s];
    };
    auto self_pos = [&](const Ctr &C) -> int { return C.live ? inv_pos[n_edges + C.c] : 0; };
    Ctr C = load_centre(tr.first, n_centers, centers, offs, species, x, dc);
    double gs = grad_of(C);
    int ips = self_pos(C);
    int ipe = s < C.N ? inv_pos[C.beg + s] : 0;  // record position of lane s's edge
    Nbr B = load_nbr(load_index(C, s, nl), species, x, wt.nl_species != 0);
    for (int c = tr.first; c < tr.hi; c += tr.step) {
        const Ctr Cn = load_centre(c + tr.step, n_centers, centers, offs, species, x, dc);
        const double gsn = grad_of(Cn);
        const int ipsn = self_pos(Cn);
        // a skipped centre (flagged request) writes only its zero self record
        if (C.work && s < width) L.g[s] = gs;
        // FP32-bulk: the g2 rows' weights per species block, from the lanes
        // that hold them (past the triplet weights in L.w)
        float *gw = reinterpret_cast<float *>(L.w) + kWEndAdj;
        if constexpr (P == 1) {
            const int slot = wp.slots[s & 15];
            const double gv = __shfl(gs, slot < 0 ? 0 : slot);
            if (s < 8 * NSB) gw[s] = slot >= 0 ? (float)gv : 0.0f;
        }
        __syncthreads();

        // ---- pairs: radial part of dE/d r_j, the accumulator's start value ----
        bool valid = false;
        double d0 = 0, d1 = 0, d2 = 0, r2 = 0, r = 0;
        const int sj = B.sj;
        if (s < C.N) {
            d0 = B.p0 - C.x0;
            d1 = B.p1 - C.x1;
            d2 = B.p2 - C.x2;
            r2 = dist2(d0, d1, d2);
            r = sqrt(r2);
            valid = r < S.rc[pidx(C.si, sj, ns)];  // r > rc is skipped; r == rc adds exactly 0
            if (!valid) {
                double *ej = edge + 3 * (size_t)ipe;
                ej[0] = 0.0;
                ej[1] = 0.0;
                ej[2] = 0.0;
            }
        }
        const unsigned long long m = __ballot(valid);
        const int M = __popcll(m);
        // two species blocks: species-0 neighbours first, list order within each
        const unsigned long long m1 = NSB == 2 ? __ballot(valid && sj == 1) : 0ull;
        const int n1 = NSB == 2 ? __popcll(m1) : 0;
        if (valid) {
            const int pos = NSB == 2 ? (sj ? (M - n1) + below(m1, s) : below(m & ~m1, s))
                                     : below(m, s);
            const int k = pidx(C.si, sj, ns);
            T fc, dfc;
            cut_pair(r, S.rc[k], S.rcinv[k], S.rcs[k], fc, dfc, R);
            double sum = 0.0;
            if constexpr (P == 1)
                sum = (double)pair_rows_f<NSB, true, false>(wp, gw, (float)r, fc, dfc, sj, nullptr);
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                if (P == 0 && q < wt.nr2) {  // L.g[slot] needs a row of the set
                    const int slot = wp.slots[(NSB == 1 ? 0 : sj) * 8 + q];
                    const double dr = r - wp.rs2[q];
                    const T ex = dexp(-wp.eta2[q] * dr * dr, R);
                    const T dphi = ex * (T)(-2.0 * wp.eta2[q] * dr) * fc + ex * dfc;
                    sum += L.g[slot] * (double)dphi;
                }
            }
            // 1/r: FP32 (one Newton step) in the FP32-bulk form
            const double ir = P == 1 ? (double)rcp_refined((float)r) : 1.0 / r;
            if constexpr (P == 1) {
                double2 *D = reinterpret_cast<double2 *>(L.px);
                float4 *V = reinterpret_cast<float4 *>(D + 2 * kW);
                float2 *W = reinterpret_cast<float2 *>(V + kW);
                const float r2f = (float)r2;
                D[pos] = make_double2(d0, d1);
                D[kW + pos] = make_double2(d2, r);
                // the species (one block) or, with the blocks' sorted slots, r in FP32
                V[pos] = make_float4((float)ir, fc, dfc, NPB > 1 ? (float)r : __int_as_float(sj));
                W[pos] = make_float2(__builtin_amdgcn_exp2f(-F.etal[0] * r2f),
                                     __builtin_amdgcn_exp2f(-F.etal[1] * r2f));
            } else {
                L.px[pos] = B.p0;
                L.py[pos] = B.p1;
                L.pz[pos] = B.p2;
                L.r[pos] = r;
                L.rinv[pos] = ir;
                L.fc[pos] = (double)fc;
                L.dfc[pos] = (double)dfc;
                L.sp[pos] = sj;
            }
            L.ip[pos] = ipe;
            L.fx[pos] = sum * d0 * ir;
            L.fy[pos] = sum * d1 * ir;
            L.fz[pos] = sum * d2 * ir;
        }
        const int jn = load_index(Cn, s, nl);
        const int ipen = s < Cn.N ? inv_pos[Cn.beg + s] : 0;
        __syncthreads();

        // ---- triplets: the active pairs of every block, lanes split over the blocks ----
        Nbr Bn;
        if (wt.has3) {
            int cnt[3], base[3];
            if (wt.rc_uniform)
                enumerate_pairs<P, NPB, true>(S, L, F, ns, M, n1, s, cnt, base);
            else
                enumerate_pairs<P, NPB, false>(S, L, F, ns, M, n1, s, cnt, base);
            Bn = load_nbr(jn, species, x, wt.nl_species != 0);
            // row weights of every block: dE/dzeta 2^(1-zeta), and times zeta lambda,
            // kept in the compute type (a loop-invariant copy costs half the registers)
            if (s < 16 * NPB) {
                const int k = s & 15, z = k >> 2, l = (k >> 1) & 1;
                const int slot = wp.slots[16 + s];
                const double ww = (C.work && slot >= 0) ? L.g[slot] * wp.coef[z] : 0.0;
                reinterpret_cast<T *>(L.w)[s] = (T)ww;
                reinterpret_cast<T *>(L.wv)[s] = (T)(ww * wp.zeta[z] * wp.lam[l]);
            }
            __syncthreads();
            const int nall = cnt[0] + (NPB > 1 ? cnt[1] + cnt[2] : 0);
            int passes = (nall + kW - 1) / kW;
            // lanes dealt to the blocks in proportion to their active pairs:
            // the fewest passes I with sum_p ceil(cnt_p / I) <= 64 lanes; lane
            // s serves block p with stride lanes(p)
            int ln[3] = {kW, 0, 0};
            if (NPB > 1 && nall > 0) {
                auto lanes_for = [&](int I) {
                    ln[0] = ceil_div_small(cnt[0], I);
                    ln[1] = ceil_div_small(cnt[1], I);
                    ln[2] = ceil_div_small(cnt[2], I);
                    return ln[0] + ln[1] + ln[2];
                };
                while (lanes_for(passes) > kW) ++passes;
            }
            const int pl = NPB == 1 ? 0 : (s < ln[0] ? 0 : (s < ln[0] + ln[1] ? 1 : 2));
            const int u = NPB == 1 ? s : s - (pl == 0 ? 0 : (pl == 1 ? ln[0] : ln[0] + ln[1]));
            const int stride = NPB == 1 ? kW : (pl == 0 ? ln[0] : (pl == 1 ? ln[1] : ln[2]));
            const bool mine = u < stride;
            const int nb = NPB == 1 ? cnt[0] : (pl == 0 ? cnt[0] : (pl == 1 ? cnt[1] : cnt[2]));
            const unsigned short *lst =
                L.l0 + (NPB == 1 ? 0 : (pl == 0 ? base[0] : (pl == 1 ? base[1] : base[2])));
            const T *w = reinterpret_cast<const T *>(L.w) + kWBlkAdj * pl;
            const T *wv = reinterpret_cast<const T *>(L.wv) + kWBlkAdj * pl;
            // this lane's species-pair cutoff (FP32-bulk)
            const int kp = blk_pair(pl);
            const float rcl = P == 1 ? (float)S.rc[kp] : 0.0f;
            const double rc2l = P == 1 ? S.rc2[kp] : 0.0;
            const float rcsl = P == 1 ? (float)S.rcs[kp] : 0.0f;
            double accd[16];  // (the fused kernel's row sums; unused here)
            for (int i = 0; i < passes; ++i) {
                const int t = u + i * stride;
                if (mine && t < nb) {
                    const int code = lst[t];
                    const int a = code & 63, b = (code >> 6) & 63;
                    if constexpr (P == 1) {
                        adj_triplet_f<NPB, false, ZG>(S, F, L, ns, rcl, rc2l, rcsl, 0, a, b, w, wv, accd);
                    } else {
                        Trip<P> q;
                        triplet<P, true>(S, wp, wt, L, ns, a, b, q);
                        const T dfa = (T)L.dfc[a], dfb = (T)L.dfc[b];
                        const T fc3 = q.fa * q.fb * q.fjk;
                        T U[2] = {(T)0, (T)0}, V[2] = {(T)0, (T)0};
#pragma unroll
                        for (int l = 0; l < 2; ++l) {
                            if (l < wt.nl) {
                                T pm[4];
                                if constexpr (P == 0 && ZG != 0)
                                    grid_powers64_m1<ZG>(q.B[l], pm);
                                else
                                    powers_m1<P>(q.B[l], nsq, wp, wt.nz, pm);
#pragma unroll
                                for (int z = 0; z < 4; ++z) {
                                    if (z < wt.nz) {
                                        const T pm1 = pm[z];
                                        const T pw = pm1 * q.B[l];
#pragma unroll
                                        for (int e = 0; e < 2; ++e) {
                                            U[e] += w[(z * 2 + l) * 2 + e] * pw;
                                            V[e] += wv[(z * 2 + l) * 2 + e] * pm1;
                                        }
                                    }
                                }
                            }
                        }
                        T SA = (T)0, SB = (T)0, SC = (T)0;
#pragma unroll
                        for (int e = 0; e < 2; ++e) {
                            if (e < wt.ne) {
                                SA += V[e] * q.E[e];
                                const T ue = U[e] * q.E[e];
                                SB += ue;
                                SC += (T)wp.eta3[e] * ue;
                            }
                        }
                        // d cos(theta) / d r_ij, d r_ik, d r_jk
                        const T dcij = (T)q.nij * q.inv2 * (T)q.ria;
                        const T dcik = (T)q.nik * q.inv2 * (T)q.rib;
                        const T dcjk = -(T)q.rjk * (T)2 * q.inv2;
                        const T Aij = fc3 * (dcij * SA - (T)(2 * q.ra) * SC) + q.fb * q.fjk * dfa * SB;
                        const T Aik = fc3 * (dcik * SA - (T)(2 * q.rb) * SC) + q.fa * q.fjk * dfb * SB;
                        const T Ajk = fc3 * (dcjk * SA - (T)(2 * q.rjk) * SC) + q.fa * q.fb * q.dfjk * SB;
                        const double a1 = (double)Aij * q.ria, a2 = (double)Aik * q.rib,
                                     a3 = (double)Ajk * q.rjkinv;
                        atomicAdd(&L.fx[a], a1 * (L.px[a] - C.x0) - a3 * q.e0);
                        atomicAdd(&L.fy[a], a1 * (L.py[a] - C.x1) - a3 * q.e1);
                        atomicAdd(&L.fz[a], a1 * (L.pz[a] - C.x2) - a3 * q.e2);
                        atomicAdd(&L.fx[b], a2 * (L.px[b] - C.x0) + a3 * q.e0);
                        atomicAdd(&L.fy[b], a2 * (L.py[b] - C.x1) + a3 * q.e1);
                        atomicAdd(&L.fz[b], a2 * (L.pz[b] - C.x2) + a3 * q.e2);
                    }
                }
            }
            __syncthreads();
        } else {
            Bn = load_nbr(jn, species, x, wt.nl_species != 0);
        }

        // ---- edge records and the centre's own record ----
        double sv[4] = {0.0, 0.0, 0.0, 0.0};
        if (s < M) {
            const double f0 = L.fx[s], f1 = L.fy[s], f2 = L.fz[s];
            double *ej = edge + 3 * (size_t)L.ip[s];
            ej[0] = f0;
            ej[1] = f1;
            ej[2] = f2;
            sv[0] = -f0;
            sv[1] = -f1;
            sv[2] = -f2;
        }
        wave_reduce<4>(sv, s);
        if (C.live && (s & 15) == 0 && (s >> 4) < 3) edge[3 * (size_t)ips + (s >> 4)] = sv[0];
        __syncthreads();
        C = Cn;
        B = Bn;
        gs = gsn;
        ips = ipsn;
        ipe = ipen;
    }
}

// centres [lo, hi) of this block for the paired kernel: the XCD slices of
// tile_range (block b on XCD b % 8), each split into contiguous runs among
// the XCD's blocks, so that consecutive centres of a run can share a wave
struct RunRange {
    int lo, hi;
};

__device__ __forceinline__ RunRange run_range(int n)
{
    const int G = gridDim.x, b = blockIdx.x;
    RunRange r;
    if (G >= 8 && (G & 7) == 0) {
        const int xcd = b & 7, local = b >> 3, per = G >> 3;
        const int chunk = (n + 7) >> 3;
        const int lo = min(n, xcd * chunk), hi = min(n, lo + chunk);
        const long len = hi - lo;
        r.lo = lo + (int)(len * local / per);
        r.hi = lo + (int)(len * (local + 1) / per);
    } else {
        r.lo = (int)((long)n * b / G);
        r.hi = (int)((long)n * (b + 1) / G);
    }
    return r;
}

// The fused forward + adjoint kernel, two centres to a wave.  Two
// consecutive centres of a block's run whose neighbour lists both have at
// most 32 entries share the pair phase: lanes 0..31 take the first centre's
// entries and lanes 32..63 the second's, and the compacted neighbours sit at
// slots 0.. and 32.. of the neighbour arrays.  Each centre's g2 row sums,
// pair enumeration, triplet passes, row reduction and records then run over
// all 64 lanes, in turn, in the same lanes and order as for a centre alone,
// so a centre's descriptors and records are bitwise the same whichever
// centre shares its wave, or none.  A centre with more than 32 entries, or
// next to one, or last in its run, takes all 64 lanes alone.  The pair phase
// is about a sixth of a centre's instructions and ran at the lane occupancy
// of the neighbour count (~26 of 64 lanes); pairing halves it.
// The triplet lanes are dealt 4-lane columns so that the row sums reduce as
// in the forward kernel; the descriptors are accumulated from the same pair
// and triplet work (the reference's gradient adds the forward into desc as
// well, src/Descriptors.cpp:165-190).
template <int P, int NSB, int NPB, int ZG, bool DC>
__global__ __launch_bounds__(64, P == 0 ? 2 : 3)
void k_fused_pair(WaveTab wt, int n_centers, const int *__restrict__ centers,
                  const int *__restrict__ offs, const int *__restrict__ nl,
                  const int *__restrict__ species, const double *__restrict__ x,
                  const double *__restrict__ dE, const int *__restrict__ inv_pos,
                  double *__restrict__ edge, int n_edges, double *__restrict__ desc, int width,
                  int maxn, DevCounts dc_in, int assign_rows)
{
    using T = typename Real<P>::T;
    const Real<P> R;
    const DevCounts dc = DC ? dc_in : DevCounts();
    extern __shared__ double lds_base[];
    const int s = threadIdx.x;
    const int ns = NSB == 2 ? 2 : wt.n_species;  // resolved: two species
    const Shared S = carve_shared(lds_base, ns);
    const Grp L = carve_group(lds_base, ns, kPair);
    stage_shared(S, wt, s);
    const WaveParams &wp = *S.par;
    const int nsq = __builtin_amdgcn_readfirstlane(wp.nsq);
    const Fast F = P == 1 ? load_fast<NPB>(S, wp) : Fast();
    if (dc.hdr) {
        n_centers = dc.hdr[0];
        n_edges = dc.hdr[2];
    }
    const RunRange rr = run_range(n_centers);
    if (rr.lo >= rr.hi) return;
    auto grad_of = [&](const Ctr &C) -> double {
        if (!C.work || s >= width) return 0.0;
        return dc.in_row ? dc.in_row[C.c][s] : dE[(size_t)C.c * width + s];
    };
    auto self_pos = [&](const Ctr &C) -> int { return C.live ? inv_pos[n_edges + C.c] : 0; };
    const bool assign = dc.out_row != nullptr || assign_rows != 0;
    auto row_in = [&](const Ctr &C) -> double {
        return (!assign && C.work && s < width) ? desc[(size_t)C.c * width + s] : 0.0;
    };
    auto row_of = [&](const Ctr &C) -> double * {
        return C.work ? (dc.out_row ? dc.out_row[C.c] : desc + (size_t)C.c * width) : nullptr;
    };
    // the unit at centre c: c alone, or c and c + 1 (returns true) when both
    // lists fit 32 lanes; past the run, centres that are not live
    auto load_unit = [&](int c, Ctr &A, Ctr &B) -> bool {
        A = load_centre(c < rr.hi ? c : n_centers, n_centers, centers, offs, species, x, dc);
        B = load_centre(c + 1 < rr.hi ? c + 1 : n_centers, n_centers, centers, offs, species, x, dc);
        return c + 1 < rr.hi && A.N <= 32 && B.N <= 32;
    };
    // lane s's list entry of a unit: its centre (the second for s >= 32 of a
    // pair) and the entry's index there
    auto entry_of = [&](bool two, const Ctr &A, const Ctr &B, int &beg, int &e) -> bool {
        const bool hb = two && s >= 32;
        e = two ? (s & 31) : s;
        beg = hb ? B.beg : A.beg;
        return e < (hb ? B.N : A.N);
    };
    Ctr A, B;
    int c = rr.lo;
    bool two = load_unit(c, A, B);
    double gA = grad_of(A), gB = two ? grad_of(B) : 0.0;
    double rA = row_in(A), rB = two ? row_in(B) : 0.0;
    int ipsA = self_pos(A), ipsB = two ? self_pos(B) : 0;
    int ipe;
    Nbr Bv;
    {
        int beg, e;
        const bool in = entry_of(two, A, B, beg, e);
        ipe = in ? inv_pos[beg + e] : 0;
        Bv = load_nbr(in ? nl[beg + e] : -1, species, x, wt.nl_species != 0);
    }
    while (c < rr.hi) {
        // the lane index, opaque to the compiler within each unit (an empty asm
        // that claims to change it): the values derived from it are formed in
        // the unit where they are used, not hoisted out of the loop and held
        // (or spilled) through all of it
        int s_unit = s;
        asm volatile("" : "+v"(s_unit));
        {
        const int s = s_unit;
        const int cn = c + (two ? 2 : 1);
        Ctr An, Bn;
        const bool twon = load_unit(cn, An, Bn);
        double *rowA = row_of(A);
        double *rowB = two ? row_of(B) : nullptr;
        if (A.work && s < width) L.g[s] = gA;
        if (two && B.work && s < width) L.gB[s] = gB;
        if (s < kWaveSlots) {
            L.d[s] = rA;
            L.dB[s] = rB;
        }
        // FP32-bulk: each centre's g2 row weights per species block (past the
        // triplet weights in L.w)
        float *gwA = reinterpret_cast<float *>(L.w) + kWEnd;
        float *gwB = gwA + 16;
        if constexpr (P == 1) {
            const int slot = wp.slots[s & 15];
            const double gvA = __shfl(gA, slot < 0 ? 0 : slot);
            const double gvB = __shfl(gB, slot < 0 ? 0 : slot);
            if (s < 8 * NSB) {
                gwA[s] = slot >= 0 ? (float)gvA : 0.0f;
                gwB[s] = slot >= 0 ? (float)gvB : 0.0f;
            }
        }
        __syncthreads();

        // ---- pairs of both centres: radial part of dE/d r_j ----
        const bool hb = two && s >= 32;
        const int e = two ? (s & 31) : s;
        const int XN = hb ? B.N : A.N, Xsi = hb ? B.si : A.si;
        bool valid = false;
        double d0 = 0, d1 = 0, d2 = 0, r2 = 0, r = 0;
        const int sj = Bv.sj;
        if (e < XN) {
            d0 = Bv.p0 - (hb ? B.x0 : A.x0);
            d1 = Bv.p1 - (hb ? B.x1 : A.x1);
            d2 = Bv.p2 - (hb ? B.x2 : A.x2);
            r2 = dist2(d0, d1, d2);
            r = sqrt(r2);
            valid = r < S.rc[pidx(Xsi, sj, ns)];  // r > rc is skipped; r == rc adds exactly 0
            if (!valid) {
                double *ej = edge + 3 * (size_t)ipe;
                ej[0] = 0.0;
                ej[1] = 0.0;
                ej[2] = 0.0;
            }
        }
        const unsigned long long m = __ballot(valid);
        const unsigned long long mA = two ? (m & 0xffffffffull) : m, mB = m & ~mA;
        const int MA = __popcll(mA), MB = __popcll(mB);
        // two species blocks: species-0 neighbours first, list order within each
        const unsigned long long m1 = NSB == 2 ? __ballot(valid && sj == 1) : 0ull;
        const int n1A = NSB == 2 ? __popcll(m1 & mA) : 0, n1B = NSB == 2 ? __popcll(m1 & mB) : 0;
        T *terms = reinterpret_cast<T *>(L.l0);  // the g2 terms, until the enumerations
        if (valid) {
            const unsigned long long mX = hb ? mB : mA, m1X = m1 & mX;
            const int MX = hb ? MB : MA, n1X = hb ? n1B : n1A;
            const int pos = (hb ? 32 : 0) + (NSB == 2 ? (sj ? (MX - n1X) + below(m1X, s)
                                                             : below(mX & ~m1X, s))
                                                      : below(mX, s));
            const int k = pidx(Xsi, sj, ns);
            T fc, dfc;
            cut_pair(r, S.rc[k], S.rcinv[k], S.rcs[k], fc, dfc, R);
            double sum = 0.0;
            if constexpr (P == 1)
                sum = (double)pair_rows_f<NSB, true, true>(wp, hb ? gwB : gwA, (float)r, fc, dfc,
                                                           sj, reinterpret_cast<float *>(terms) + 8 * pos);
            const double *gX = hb ? L.gB : L.g;
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                if (P == 0) {
                    double v = 0.0;
                    if (q < wt.nr2) {  // the centre's dE/dzeta row needs a row of the set
                        const int slot = wp.slots[(NSB == 1 ? 0 : sj) * 8 + q];
                        const double dr = r - wp.rs2[q];
                        const T ex = dexp(-wp.eta2[q] * dr * dr, R);
                        const T dphi = ex * (T)(-2.0 * wp.eta2[q] * dr) * fc + ex * dfc;
                        sum += gX[slot] * (double)dphi;
                        v = (double)(ex * fc);  // the g2 term, as the forward kernel forms it
                    }
                    terms[8 * pos + q] = (T)v;
                }
            }
            // 1/r: FP32 (one Newton step) in the FP32-bulk form
            const double ir = P == 1 ? (double)rcp_refined((float)r) : 1.0 / r;
            if constexpr (P == 1) {
                double2 *D = reinterpret_cast<double2 *>(L.px);
                float4 *V = reinterpret_cast<float4 *>(D + 2 * kW);
                float2 *W = reinterpret_cast<float2 *>(V + kW);
                const float r2f = (float)r2;
                D[pos] = make_double2(d0, d1);
                D[kW + pos] = make_double2(d2, r);
                V[pos] = make_float4((float)ir, fc, dfc, __int_as_float(sj));
                W[pos] = make_float2(__builtin_amdgcn_exp2f(-F.etal[0] * r2f),
                                     __builtin_amdgcn_exp2f(-F.etal[1] * r2f));
            } else {
                L.px[pos] = Bv.p0;
                L.py[pos] = Bv.p1;
                L.pz[pos] = Bv.p2;
                L.r[pos] = r;
                L.rinv[pos] = ir;
                L.fc[pos] = (double)fc;
                L.dfc[pos] = (double)dfc;
                L.sp[pos] = sj;
            }
            L.ip[pos] = ipe;
            L.fx[pos] = sum * d0 * ir;
            L.fy[pos] = sum * d1 * ir;
            L.fz[pos] = sum * d2 * ir;
        }
        // the next unit's list entries
        int jn, ipen;
        {
            int beg, en;
            const bool in = entry_of(twon, An, Bn, beg, en);
            jn = in ? nl[beg + en] : -1;
            ipen = in ? inv_pos[beg + en] : 0;
        }
        __syncthreads();
        // a centre of at most 32 compacted neighbours sums its g2 rows (and
        // its own record, below) over 32 lanes, one half-wave per centre; a
        // larger one, never paired, over all 64
        const bool halves = two || MA <= 32;
        if (halves)
            g2_rows_sum_half<NSB>(terms + (hb ? 8 * 32 : 0), hb ? MB : MA,
                                  hb ? MB - n1B : MA - n1A, s, wp.slots, hb ? L.dB : L.d,
                                  s < 32 || two);
        else
            g2_rows_sum<NSB>(terms, MA, MA - n1A, s, wp.slots, L.d);
        __syncthreads();

        // ---- triplets of each centre in turn, lanes split over its blocks ----
        Nbr Bvn;
        if (wt.has3) {
#pragma clang loop unroll(disable)
            for (int h = 0; h < (two ? 2 : 1); ++h) {
                // the lane index opaque again for each centre (as for each unit
                // above): its derived values are formed per centre
                int s_cen = s;
                asm volatile("" : "+v"(s_cen));
                {
                const int s = s_cen;
                const int off = h ? 32 : 0;
                const int M = h ? MB : MA, n1 = h ? n1B : n1A;
                const Ctr &X = h ? B : A;
                double *drow = h ? L.dB : L.d;
                const double *grow = h ? L.gB : L.g;
                int cnt[3], base[3];
                if (wt.rc_uniform)
                    enumerate_pairs<P, NPB, true>(S, L, F, ns, M, n1, s, cnt, base, off);
                else
                    enumerate_pairs<P, NPB, false>(S, L, F, ns, M, n1, s, cnt, base, off);
                if (h == 0) Bvn = load_nbr(jn, species, x, wt.nl_species != 0);
                // row weights of every block: dE/dzeta 2^(1-zeta), and times zeta lambda
                if (s < 16 * NPB) {
                    const int k = s & 15, z = k >> 2, l = (k >> 1) & 1;
                    const int slot = wp.slots[16 + s];
                    const double ww = (X.work && slot >= 0) ? grow[slot] * wp.coef[z] : 0.0;
                    const int at = kWBlk * (s >> 4) + k;
                    reinterpret_cast<T *>(L.w)[at] = (T)ww;
                    reinterpret_cast<T *>(L.wv)[at] = (T)(ww * wp.zeta[z] * wp.lam[l]);
                }
                __syncthreads();
                // columns of 4 lanes (s, s ^ 16, s ^ 32, s ^ 48; column s & 15)
                // dealt to the blocks in proportion to their active pairs, the
                // fewest passes I with sum_p ceil(cnt_p / 4I) <= 16 columns
                const int nall = cnt[0] + (NPB > 1 ? cnt[1] + cnt[2] : 0);
                int passes = (nall + kW - 1) / kW;
                int nc[3] = {16, 0, 0};
                const int col = s & 15, rowq = s >> 4;
                if (NPB > 1 && nall > 0) {
                    const int c40 = (cnt[0] + 3) >> 2, c41 = (cnt[1] + 3) >> 2, c42 = (cnt[2] + 3) >> 2;
                    auto cols_for = [&](int I) {
                        nc[0] = ceil_div_small(c40, I);
                        nc[1] = ceil_div_small(c41, I);
                        nc[2] = ceil_div_small(c42, I);
                        return nc[0] + nc[1] + nc[2];
                    };
                    while (cols_for(passes) > 16) ++passes;
                }
                const int pl = NPB == 1 ? 0 : (col < nc[0] ? 0 : (col < nc[0] + nc[1] ? 1 : 2));
                const int cs = pl == 0 ? 0 : (pl == 1 ? nc[0] : nc[0] + nc[1]);
                const int ncol = pl == 0 ? nc[0] : (pl == 1 ? nc[1] : nc[2]);
                const int u = NPB == 1 ? s : (col - cs) + rowq * ncol;
                const int stride = NPB == 1 ? kW : 4 * ncol;
                const bool mine = NPB == 1 || col < cs + ncol;
                const int nb = NPB == 1 ? cnt[0] : (pl == 0 ? cnt[0] : (pl == 1 ? cnt[1] : cnt[2]));
                const unsigned short *lst =
                    L.l0 + (NPB == 1 ? 0 : (pl == 0 ? base[0] : (pl == 1 ? base[1] : base[2])));
                const T *w = reinterpret_cast<const T *>(L.w) + kWBlk * pl;
                const T *wv = reinterpret_cast<const T *>(L.wv) + kWBlk * pl;
                const int kp = blk_pair(pl);
                const float rcl = P == 1 ? (float)S.rc[kp] : 0.0f;
                const double rc2l = P == 1 ? S.rc2[kp] : 0.0;
                const float rcsl = P == 1 ? (float)S.rcs[kp] : 0.0f;
                double accd[16];
#pragma unroll
                for (int k = 0; k < 16; ++k) accd[k] = 0.0;
                for (int i = 0; i < passes; ++i) {
                    const int t = u + i * stride;
                    if (mine && t < nb) {
                        const int code = lst[t];
                        // slots relative to the centre's first (off)
                        const int ra = code & 63, rb = (code >> 6) & 63;
                        if constexpr (P == 1) {
                            adj_triplet_f<NPB, true, ZG>(S, F, L, ns, rcl, rc2l, rcsl, off, ra, rb, w, wv,
                                                         accd);
                        } else {
                            const int a = ra + off, b = rb + off;
                            Trip<P> q;
                            triplet<P, true>(S, wp, wt, L, ns, a, b, q);
                            const T dfa = (T)L.dfc[a], dfb = (T)L.dfc[b];
                            const T fc3 = q.fa * q.fb * q.fjk;
                            double te[2];
#pragma unroll
                            for (int ee = 0; ee < 2; ++ee) te[ee] = (double)(q.E[ee] * fc3);
                            T U[2] = {(T)0, (T)0}, V[2] = {(T)0, (T)0};
#pragma unroll
                            for (int l = 0; l < 2; ++l) {
                                if (l < wt.nl) {
                                    T pm[4];
                                    if constexpr (ZG != 0)
                                        grid_powers64_m1<ZG>(q.B[l], pm);
                                    else
                                        powers_m1<P>(q.B[l], nsq, wp, wt.nz, pm);
#pragma unroll
                                    for (int z = 0; z < 4; ++z) {
                                        if (z < wt.nz) {
                                            const T pm1 = pm[z];
                                            const T pw = pm1 * q.B[l];
#pragma unroll
                                            for (int ee = 0; ee < 2; ++ee) {
                                                U[ee] += w[(z * 2 + l) * 2 + ee] * pw;
                                                V[ee] += wv[(z * 2 + l) * 2 + ee] * pm1;
                                            }
                                            const double pz = (double)pw;  // 2^(1 - zeta) at the row sum
#pragma unroll
                                            for (int ee = 0; ee < 2; ++ee)
                                                if (ee < wt.ne)
                                                    accd[(z * 2 + l) * 2 + ee] =
                                                        fma(pz, te[ee], accd[(z * 2 + l) * 2 + ee]);
                                        }
                                    }
                                }
                            }
                            T SA = (T)0, SB = (T)0, SC = (T)0;
#pragma unroll
                            for (int ee = 0; ee < 2; ++ee) {
                                if (ee < wt.ne) {
                                    SA += V[ee] * q.E[ee];
                                    const T ue = U[ee] * q.E[ee];
                                    SB += ue;
                                    SC += (T)wp.eta3[ee] * ue;
                                }
                            }
                            const T dcij = (T)q.nij * q.inv2 * (T)q.ria;
                            const T dcik = (T)q.nik * q.inv2 * (T)q.rib;
                            const T dcjk = -(T)q.rjk * (T)2 * q.inv2;
                            const T Aij = fc3 * (dcij * SA - (T)(2 * q.ra) * SC) + q.fb * q.fjk * dfa * SB;
                            const T Aik = fc3 * (dcik * SA - (T)(2 * q.rb) * SC) + q.fa * q.fjk * dfb * SB;
                            const T Ajk = fc3 * (dcjk * SA - (T)(2 * q.rjk) * SC) + q.fa * q.fb * q.dfjk * SB;
                            const double a1 = (double)Aij * q.ria, a2 = (double)Aik * q.rib,
                                         a3 = (double)Ajk * q.rjkinv;
                            atomicAdd(&L.fx[a], a1 * (L.px[a] - X.x0) - a3 * q.e0);
                            atomicAdd(&L.fy[a], a1 * (L.py[a] - X.x1) - a3 * q.e1);
                            atomicAdd(&L.fz[a], a1 * (L.pz[a] - X.x2) - a3 * q.e2);
                            atomicAdd(&L.fx[b], a2 * (L.px[b] - X.x0) + a3 * q.e0);
                            atomicAdd(&L.fy[b], a2 * (L.py[b] - X.x1) + a3 * q.e1);
                            atomicAdd(&L.fz[b], a2 * (L.pz[b] - X.x2) + a3 * q.e2);
                        }
                    }
                }
                __syncthreads();
                // the row sums: a column's 4 lanes with two register swaps, then
                // the columns of a block through LDS (the pair list's region,
                // now read), in order, all in FP64; lane s keeps values
                // 4 (s >> 4) .. +3
                double v4[4];
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    double a = accd[q], b = accd[q + 8];
                    xswap<32>(a, b);
                    accd[q] = a + b;
                }
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    double a = accd[q], b = accd[q + 4];
                    xswap<16>(a, b);
                    accd[q] = a + b;
                }
#pragma unroll
                for (int q = 0; q < 4; ++q) v4[q] = accd[q];
                double *buf = reinterpret_cast<double *>(L.l0);
#pragma unroll
                for (int q = 0; q < 4; ++q) buf[kColStr * (4 * rowq + q) + col] = v4[q];
                __syncthreads();
                if (rowq < NPB) {
                    const int c0 = rowq == 0 ? 0 : (rowq == 1 ? nc[0] : nc[0] + nc[1]);
                    const int c1 = c0 + (rowq == 0 ? nc[0] : (rowq == 1 ? nc[1] : nc[2]));
                    double sum = 0.0;
                    for (int cc = c0; cc < c1; ++cc) sum += buf[kColStr * col + cc];
                    const int slot = wp.slots[16 + s];
                    // each row takes its 2^(1 - zeta) here (a power of two for the integer
                    // zeta of the grids: the same bits as scaling every term)
                    sum *= wp.coef[col >> 2];
                    if (c1 > c0 && slot >= 0) drow[slot] += sum;
                }
                // the column sums are read before the next enumeration reuses the region
                __syncthreads();
                }
            }
        } else {
            Bvn = load_nbr(jn, species, x, wt.nl_species != 0);
        }
        // the next unit's rows, loaded here rather than a unit ahead: held
        // through the triplet passes they would cost registers there
        const double gAn = grad_of(An), gBn = twon ? grad_of(Bn) : 0.0;
        const double rAn = row_in(An), rBn = twon ? row_in(Bn) : 0.0;
        const int ipsAn = self_pos(An), ipsBn = twon ? self_pos(Bn) : 0;

        // ---- edge records and each centre's own record ----
        // (lane s holds slot s: the first centre's slots 0.., the second's 32..)
        {
            double sv[4] = {0.0, 0.0, 0.0, 0.0};
            if (s < 32 ? s < MA : (halves ? (two && s - 32 < MB) : s < MA)) {
                const double f0 = L.fx[s], f1 = L.fy[s], f2 = L.fz[s];
                double *ej = edge + 3 * (size_t)L.ip[s];
                ej[0] = f0;
                ej[1] = f1;
                ej[2] = f2;
                sv[0] = -f0;
                sv[1] = -f1;
                sv[2] = -f2;
            }
            if (halves) {
                Halve<4, 2, kW / 4>::run(sv, s);  // each half-wave's total of value (s >> 3) & 3
                const int v = (s >> 3) & 3;
                if ((s & 7) == 0 && v < 3 && (s < 32 || two))
                    edge[3 * (size_t)(s < 32 ? ipsA : ipsB) + v] = sv[0];
            } else {
                wave_reduce<4>(sv, s);
                if ((s & 15) == 0 && (s >> 4) < 3) edge[3 * (size_t)ipsA + (s >> 4)] = sv[0];
            }
        }
        __syncthreads();
        if (rowA && s < width) rowA[s] = L.d[s];
        if (rowB && s < width) rowB[s] = L.dB[s];
        c = cn;
        A = An;
        B = Bn;
        two = twon;
        gA = gAn;
        gB = gBn;
        rA = rAn;
        rB = rBn;
        ipsA = ipsAn;
        ipsB = ipsBn;
        ipe = ipen;
        Bv = Bvn;
        }
    }
}

int grid_for(int n_centers, int cap)
{
    int g = std::min(n_centers, cap > 0 ? cap : 16384);
    if (g >= 8) g = (g + 7) & ~7;
    return std::max(g, 1);
}

size_t lds_for(int max_nn, int kind, int ns, int precision)
{
    const int maxn = std::max(max_nn, 1);
    return sizeof(double) * shared_doubles(ns) +
           group_bytes(maxn, kind, precision == 0 ? (int)sizeof(double) : (int)sizeof(float));
}

}  // namespace

#ifndef LDSF_WAVE_FP64_UNIT
int wave_group_size(int max_nn, int n_species)
{
    // one wavefront per centre, one lane per neighbour-list entry
    if (max_nn > kW) return 0;
    return lds_for(max_nn, kPair, n_species, 0) <= kWaveMaxLds ? kW : 0;
}

size_t wave_lds_bytes(int max_nn, bool adjoint, int n_species, int group)
{
    (void)group;
    return lds_for(max_nn, adjoint ? kAdj : kFwd, n_species, 0);
}

bool wave_build(const ldsf_sf *sf, WaveTab *wt, WaveParams *wp)
{
    std::memset(wt, 0, sizeof(*wt));
    std::memset(wp, 0, sizeof(*wp));
    int *slots = wp->slots;
    for (int k = 0; k < kWaveSlots; ++k) slots[k] = -1;
    const int ns = (int)sf->species.size();
    wt->n_species = ns;
    if (ns > kMaxWaveSpecies) return false;
    if (sf->species_resolved && ns > 2) return false;
    wt->rc_uniform = 1;
    for (double v : sf->rcut) wt->rc_uniform &= v == sf->rcut[0] ? 1 : 0;
    wt->nsb = (sf->species_resolved && ns == 2) ? 2 : 1;
    // the FP32-bulk kernels give each two-species block one cutoff, rc(0,1)
    // for the mixed block whichever of its neighbours comes first in the list
    if (wt->nsb == 2 && sf->rcut[1] != sf->rcut[2]) return false;
    wt->npb = (sf->species_resolved && ns == 2) ? 3 : 1;
    // two-body rows: g2 only, at most 8
    const int n2 = (int)sf->k2.size();
    if (n2 > 8) return false;
    for (int q = 0; q < n2; ++q) {
        if (sf->k2[q] != G2) return false;
        wp->eta2[q] = sf->a2[q];
        wp->rs2[q] = sf->b2[q];
        wp->e2l[q] = (float)(sf->a2[q] * 1.4426950408889634);
        wp->m2e[q] = (float)(-2.0 * sf->a2[q]);
        wp->rs2f[q] = (float)sf->b2[q];
        for (int s = 0; s < wt->nsb; ++s) slots[s * 8 + q] = sf->slot2[q] + s * sf->str2[q];
    }
    wt->nr2 = n2;
    // FP32-bulk: the wave kernels form each term's exponent x in FP32, with a
    // relative error of about 3e-7 |x| (the roundings of eta log2(e), r - Rs,
    // its square and the product); every x over the cutoff sphere must stay
    // within kFp32ExpMax, where that is at most 1e-5 (and far from FP32's
    // underflow at 87).  The largest g2 exponent over r in [0, rc] is at an
    // end of the interval; the g4 factors' product reaches eta 3 rc^2.
    double rcmax = 0.0;
    for (double v : sf->rcut) rcmax = std::max(rcmax, v);
    wt->fp32_exp_ok = 1;
    for (int q = 0; q < n2; ++q) {
        const double eta = sf->a2[q], rs = sf->b2[q];
        const double xm = std::max(eta * rs * rs, eta * (rcmax - rs) * (rcmax - rs));
        if (!(eta >= 0.0) || !(xm <= kFp32ExpMax)) wt->fp32_exp_ok = 0;
    }
    for (size_t q = 0; q < sf->k3.size(); ++q) {
        const double eta = sf->eta3[q];
        if (!(eta >= 0.0) || !(3.0 * eta * rcmax * rcmax <= kFp32ExpMax)) wt->fp32_exp_ok = 0;
    }
    for (int l = 0; l < 2; ++l) {  // neutral lambda = 0 until the grid sets it
        wp->lp[l] = 0.5;
        wp->lm[l] = 0.5;
    }
    // three-body rows: g4 forming a zeta x lambda x eta grid
    const int n3 = (int)sf->k3.size();
    if (n3 == 0) return true;
    // the distinct zeta, lambda and eta of the rows (no heap: the allocation-
    // free entry point builds these tables on every call)
    double zs[5], ls[3], es[3];
    int nzs = 0, nls = 0, nes = 0;
    auto idx = [](double *v, int &n, int cap, double x) {
        for (int k = 0; k < n; ++k)
            if (v[k] == x) return k;
        if (n == cap) return -1;
        v[n] = x;
        return n++;
    };
    int seen[16];  // row of grid point (z * 2 + l) * 2 + e, or -1
    for (int k = 0; k < 16; ++k) seen[k] = -1;
    int n_seen = 0;
    for (int q = 0; q < n3; ++q) {
        if (sf->k3[q] != G4) return false;
        const int z = idx(zs, nzs, 4, sf->zeta3[q]), l = idx(ls, nls, 2, sf->lam3[q]),
                  e = idx(es, nes, 2, sf->eta3[q]);
        if (z < 0 || l < 0 || e < 0) return false;
        const int k = (z * 2 + l) * 2 + e;
        if (seen[k] >= 0) return false;
        seen[k] = q;
        ++n_seen;
    }
    if (n_seen != nzs * nls * nes) return false;
    // the adjoint stages a centre's whole descriptor row (<= kWaveSlots entries)
    if (n2 * wt->nsb + n3 * wt->npb > kWaveSlots) return false;
    wt->has3 = 1;
    wt->nz = nzs;
    wt->nl = nls;
    wt->ne = nes;
    // the power-of-two grids with compile-time squaring chains (ZGrid)
    if (nzs == 4 && zs[0] == 1.0 && zs[1] == 2.0 && zs[2] == 4.0)
        wt->zgrid = zs[3] == 16.0 ? 1 : (zs[3] == 8.0 ? 2 : 0);
    for (int z = 0; z < wt->nz; ++z) {
        wp->zeta[z] = zs[z];
        wp->coef[z] = std::pow(2, 1 - zs[z]);
        int iz = -1;
        if (zs[z] >= 0 && zs[z] <= 63 && zs[z] == std::floor(zs[z])) iz = (int)zs[z];
        wp->izeta[z] = iz;
        for (int k = 5; k >= 1; --k)
            if (iz >= (1 << k)) { wp->nsq = std::max(wp->nsq, k); break; }
    }
    for (int l = 0; l < wt->nl; ++l) {
        wp->lam[l] = ls[l];
        wp->lamcode[l] = ls[l] == 1.0 ? 1 : (ls[l] == -1.0 ? -1 : 0);
    }
    // entries past the grid stay neutral (lambda 0, zeta 0, coefficient 0, eta
    // 0): the FP32-bulk kernels evaluate the full 4 x 2 x 2 grid without
    // branches and those rows have no slot
    for (int l = 0; l < 2; ++l) {
        wp->lp[l] = 0.5 * (1.0 + wp->lam[l]);
        wp->lm[l] = 0.5 * (1.0 - wp->lam[l]);
    }
    for (int e = 0; e < wt->ne; ++e) wp->eta3[e] = es[e];
    for (int k = 0; k < 16; ++k) {
        const int q = seen[k];
        if (q < 0) continue;
        for (int p = 0; p < wt->npb; ++p) slots[16 + p * 16 + k] = sf->slot3[q] + p * sf->str3[q];
    }
    return true;
}

#endif  // LDSF_WAVE_FP64_UNIT

namespace {

// the launches of one precision form P (the FP64 instances are compiled in
// units of their own, for their compiler options: sf_wave_fp64.hip, the
// forward and adjoint kernels, and sf_wave_fp64_fused.hip, the fused kernel)
template <int P>
hipError_t launch_fwd_form(const WaveTab &wt, int n_centers, const int *centers, const int *offs,
                           const int *nl, const int *species, const double *x, double *desc,
                           int width, int max_nn, int group, hipStream_t s, DevCounts dc)
{
    if (n_centers <= 0) return hipSuccess;
    if (group != kW || max_nn > kW) return hipErrorInvalidValue;
    const int maxn = std::max(max_nn, 1);
    const size_t lds = lds_for(max_nn, kFwd, wt.n_species, P);
    const dim3 grid(grid_for(n_centers, dc.max_blocks)), block(kW);
    const bool use_dc = dc.hdr || dc.req_flag || dc.out_row || dc.in_row;
#define LDSF_FWD_DC(NSB, NPB, ZG, DCV)                                                        \
    hipLaunchKernelGGL((k_fwd_wave<P, NSB, NPB, ZG, DCV>), grid, block, lds, s, wt, n_centers, \
                       centers, offs, nl, species, x, desc, width, maxn, dc)
#define LDSF_FWD(NSB, NPB, ZG)                        \
    do {                                              \
        if (use_dc) LDSF_FWD_DC(NSB, NPB, ZG, true);  \
        else LDSF_FWD_DC(NSB, NPB, ZG, false);        \
    } while (0)
#define LDSF_FWD_P(NSB, NPB)                           \
    do {                                               \
        if (wt.zgrid == 1) LDSF_FWD(NSB, NPB, 1);      \
        else if (wt.zgrid == 2) LDSF_FWD(NSB, NPB, 2); \
        else LDSF_FWD(NSB, NPB, 0);                    \
    } while (0)
    if (wt.nsb == 1) LDSF_FWD_P(1, 1); else LDSF_FWD_P(2, 3);
#undef LDSF_FWD_P
#undef LDSF_FWD
#undef LDSF_FWD_DC
    return hipGetLastError();
}

// the adjoint (FUSED false) or the fused forward + adjoint (FUSED true)
template <int P, bool FUSED>
hipError_t launch_adj_form(const WaveTab &wt, int n_centers, const int *centers, const int *offs,
                           const int *nl, const int *species, const double *x, const double *dE,
                           const int *inv_pos, double *edge, int n_edges, int width, int max_nn,
                           int group, hipStream_t s, DevCounts dc, double *desc, bool assign)
{
    if (n_centers <= 0) return hipSuccess;
    if (group != kW || max_nn > kW) return hipErrorInvalidValue;
    const int maxn = std::max(max_nn, 1);
    const size_t lds = lds_for(max_nn, FUSED ? kPair : kAdj, wt.n_species, P);
    const dim3 grid(grid_for(n_centers, dc.max_blocks)), block(kW);
    const bool use_dc = dc.hdr || dc.req_flag || dc.out_row || dc.in_row;
#define LDSF_ADJ_DC(NSB, NPB, ZG, DCV)                                                            \
    hipLaunchKernelGGL((k_adj_wave<P, NSB, NPB, ZG, DCV>), grid, block, lds, s, wt, n_centers,    \
                       centers, offs, nl, species, x, dE, inv_pos, edge, n_edges, width, maxn, dc)
#define LDSF_PAIR_DC(NSB, NPB, ZG, DCV)                                                          \
    hipLaunchKernelGGL((k_fused_pair<P, NSB, NPB, ZG, DCV>), grid, block, lds, s, wt, n_centers, \
                       centers, offs, nl, species, x, dE, inv_pos, edge, n_edges, desc, width,   \
                       maxn, dc, assign ? 1 : 0)
#define LDSF_ADJ(NSB, NPB, ZG)                                          \
    do {                                                                \
        if constexpr (FUSED) {                                          \
            if (use_dc) LDSF_PAIR_DC(NSB, NPB, ZG, true);               \
            else LDSF_PAIR_DC(NSB, NPB, ZG, false);                     \
        } else {                                                        \
            if (use_dc) LDSF_ADJ_DC(NSB, NPB, ZG, true);                \
            else LDSF_ADJ_DC(NSB, NPB, ZG, false);                      \
        }                                                               \
    } while (0)
#define LDSF_ADJ_P(NSB, NPB)                           \
    do {                                               \
        if (wt.zgrid == 1) LDSF_ADJ(NSB, NPB, 1);      \
        else if (wt.zgrid == 2) LDSF_ADJ(NSB, NPB, 2); \
        else LDSF_ADJ(NSB, NPB, 0);                    \
    } while (0)
    if (wt.nsb == 1) LDSF_ADJ_P(1, 1); else LDSF_ADJ_P(2, 3);
#undef LDSF_ADJ_P
#undef LDSF_ADJ
#undef LDSF_PAIR_DC
#undef LDSF_ADJ_DC
    return hipGetLastError();
}

}  // namespace

#if !defined(LDSF_WAVE_FP64_UNIT)
hipError_t launch_fwd_wave(const WaveTab &wt, int precision, int n_centers, const int *centers,
                           const int *offs, const int *nl, const int *species, const double *x,
                           double *desc, int width, int max_nn, int group, hipStream_t s,
                           DevCounts dc)
{
    if (precision == 0)
        return launch_fwd_wave_fp64(wt, n_centers, centers, offs, nl, species, x, desc, width,
                                    max_nn, group, s, dc);
    return launch_fwd_form<1>(wt, n_centers, centers, offs, nl, species, x, desc, width, max_nn,
                              group, s, dc);
}

hipError_t launch_adj_wave(const WaveTab &wt, int precision, int n_centers, const int *centers,
                           const int *offs, const int *nl, const int *species, const double *x,
                           const double *dE, const int *inv_pos, double *edge, int n_edges,
                           int width, int max_nn, int group, hipStream_t s, DevCounts dc,
                           double *desc, bool assign)
{
    const bool fused = desc != nullptr || dc.out_row != nullptr;
    if (precision == 0)
        return (fused ? launch_fused_wave_fp64 : launch_adj_wave_fp64)(
            wt, n_centers, centers, offs, nl, species, x, dE, inv_pos, edge, n_edges, width,
            max_nn, group, s, dc, desc, assign);
    return (fused ? launch_adj_form<1, true> : launch_adj_form<1, false>)(
        wt, n_centers, centers, offs, nl, species, x, dE, inv_pos, edge, n_edges, width, max_nn,
        group, s, dc, desc, assign);
}
#elif LDSF_WAVE_FP64_UNIT == 1
hipError_t launch_fwd_wave_fp64(const WaveTab &wt, int n_centers, const int *centers,
                                const int *offs, const int *nl, const int *species, const double *x,
                                double *desc, int width, int max_nn, int group, hipStream_t s,
                                DevCounts dc)
{
    return launch_fwd_form<0>(wt, n_centers, centers, offs, nl, species, x, desc, width, max_nn,
                              group, s, dc);
}

hipError_t launch_adj_wave_fp64(const WaveTab &wt, int n_centers, const int *centers,
                                const int *offs, const int *nl, const int *species, const double *x,
                                const double *dE, const int *inv_pos, double *edge, int n_edges,
                                int width, int max_nn, int group, hipStream_t s, DevCounts dc,
                                double *desc, bool assign)
{
    return launch_adj_form<0, false>(wt, n_centers, centers, offs, nl, species, x, dE, inv_pos,
                                     edge, n_edges, width, max_nn, group, s, dc, desc, assign);
}
#elif LDSF_WAVE_FP64_UNIT == 2
hipError_t launch_fused_wave_fp64(const WaveTab &wt, int n_centers, const int *centers,
                                  const int *offs, const int *nl, const int *species,
                                  const double *x, const double *dE, const int *inv_pos,
                                  double *edge, int n_edges, int width, int max_nn, int group,
                                  hipStream_t s, DevCounts dc, double *desc, bool assign)
{
    return launch_adj_form<0, true>(wt, n_centers, centers, offs, nl, species, x, dE, inv_pos,
                                    edge, n_edges, width, max_nn, group, s, dc, desc, assign);
}
#endif  // LDSF_WAVE_FP64_UNIT

}  // namespace ldsf
