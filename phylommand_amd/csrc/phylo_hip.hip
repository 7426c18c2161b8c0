// phylo_hip.hip -- libphylo_hip: Felsenstein pruning on MI355X (gfx950).
//
// Design
//  * A likelihood evaluation is compiled on the host into a "pruning
//    program": every internal node once, children before parents (the
//    reference recurses left subtree, right subtree, node;
//    simpleML.cpp:127-184), each op naming where its two child partials come
//    from (a tip row, a stored partial, or one of two carries holding the
//    result of an earlier op of the same program).
//  * One launch walks the whole program.  Each work-item owns one site and
//    runs every op for it, so no inter-workgroup dependency exists and all
//    postorder levels are fused into the launch.  Only partials the walk
//    reads back from memory are stored at all.
//  * Partials are stored site-major and state-interleaved:
//    clv[slot][cat][site][4] (one 16-byte float4 per work-item per category),
//    so every global load/store of a wavefront is one contiguous 1 KiB burst.
//    DNA tips are one byte per (taxon, site) holding the 4-bit state mask.
//  * Arithmetic is FP64; storage is FP32 (default) or FP64.  P matrices are
//    FP64, indexed by the child node and category, and are wave-uniform so
//    the compiler keeps them in scalar registers.
//  * Partials are rescaled by exact powers of two (frexp / ldexp): a stored
//    one so its largest entry is in [0.5, 1), a carried FP64 one (DNA
//    walker) only when its largest entry falls below 2^-256; the integer
//    exponents add up in one scaler per (node, category, site).
//  * Per-site lnL = log(sum_c w_c sum_i rw_i L_ci 2^scaler_c) in FP64,
//    reduced per workgroup with wavefront shuffles, then summed in a fixed
//    order with compensation by a second one-wave launch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/phylo_hip.h"
#include "model_math.h"

namespace {

constexpr int kBlock = 256;
constexpr int kMaxCat = 8;
constexpr int kMaxStates = 32;
constexpr double kLn2 = 0.69314718055994530942;
constexpr int kRootCatOff = 32;   // roots[]: per-state weights [0, 32), category weights after

// Workgroups per CU below which a whole-tree launch has too few waves to
// hide its dependent chain: every site block is then walked as subtree
// pieces plus the top program (the tail split applied to all blocks), which
// multiplies the independent waves by the number of pieces.
constexpr int kSplitAllBelowPerCu = 4;

// Site groups per wave of the MFMA walker: two independent MFMA chains per
// wave share each P operand load and overlap their memory latency.
constexpr int kMfmaGroups = 2;

enum SrcKind : int { SRC_TIP = 0, SRC_CLV = 1, SRC_CARRY_A = 2, SRC_CARRY_B = 3, SRC_NONE = 4 };
// OP_SWAP: the children were swapped into kind order (order_children); the
// stochastic rounding of a stored result is keyed by them in their original
// order.
// OP_SH_*: one-hot hints for the four commonest child shapes (children in
// kind order), set by order_children; the DNA walker tests them before its
// switch, the matrix walker runs a specialised op for each.
enum OpFlags : int {
    OP_ROOT = 1, OP_TO_B = 2, OP_SWAP = 4,
    OP_SH_AT = 8, OP_SH_TT = 16, OP_SH_AB = 32, OP_SH_AC = 64,
    OP_SH_ALL = OP_SH_AT | OP_SH_TT | OP_SH_AB | OP_SH_AC
};

struct alignas(16) Op {
    int l_node, r_node;   // child node ids: P matrices are per (child node, category)
    int l_kind, r_kind;   // SrcKind
    int l_idx, r_idx;     // tip row or partial slot
    int out_slot;         // slot to store the result in, -1 for none
    int flags;            // OpFlags
};

// ---------------------------------------------------------------------------
// Device helpers
// ---------------------------------------------------------------------------

__device__ inline void load4(const float* p, double v[4]) {
    const float4 x = *reinterpret_cast<const float4*>(p);
    v[0] = x.x; v[1] = x.y; v[2] = x.z; v[3] = x.w;
}
__device__ inline void load4(const double* p, double v[4]) {
    const double2 a = reinterpret_cast<const double2*>(p)[0];
    const double2 b = reinterpret_cast<const double2*>(p)[1];
    v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y;
}
__device__ inline void load_raw4(const float* p, float v[4]) {
    const float4 x = *reinterpret_cast<const float4*>(p);
    v[0] = x.x; v[1] = x.y; v[2] = x.z; v[3] = x.w;
}
__device__ inline void load_raw4(const double* p, double v[4]) { load4(p, v); }
__device__ inline void store4(float* p, const float v[4]) {
    *reinterpret_cast<float4*>(p) = make_float4(v[0], v[1], v[2], v[3]);
}
__device__ inline void store4(double* p, const double v[4]) {
    reinterpret_cast<double2*>(p)[0] = make_double2(v[0], v[1]);
    reinterpret_cast<double2*>(p)[1] = make_double2(v[2], v[3]);
}

// 32-bit integer hash (Wellons' "lowbias32" xorshift-multiply mix).
__device__ inline uint32_t mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// Rounding of a partial (x >= 0) to the storage type.  FP32 storage rounds
// stochastically: 29 random bits are added below the float's last mantissa
// bit of the FP64 value and the bits a float does not hold are cut off, so x
// rounds up with probability equal to its distance from the float below.  A
// subtree with few tips produces the same few values at every site, so
// round-to-nearest would repeat the same error at thousands of sites and the
// total lnL would drift (+4.6e-3 at 128 taxa x 100,000 sites); with
// per-site random bits the errors average out.  The bits come from a hash of
// (site, node, category) stepped by a Weyl sequence per state, so results
// are reproducible and the same node rounds the same way in every program
// that computes it.  (Zero stays zero; values below the float normal range
// end as the nearest float of the truncated value.)
// The dithered cut and the power-of-two scaling commute (the cut reads only
// mantissa bits), so a walker may cut before or after scaling a partial.
__device__ inline double round_cut(double x, uint32_t bits, float) {
    // 64-bit add of the 29 dither bits, written on 32-bit halves
    const uint32_t lo = (uint32_t)__double2loint(x);
    const uint32_t sum = lo + (bits >> 3);
    const uint32_t hi = (uint32_t)__double2hiint(x) + (sum < lo ? 1u : 0u);
    return __hiloint2double((int)hi, (int)(sum & 0xe0000000u));   // <= 24 significant bits
}
__device__ inline double round_cut(double x, uint32_t, double) { return x; }

// Random bits for the partials of one (site, node, category); state i uses
// key + i * 0x9e3779b9.  The node is identified by its two children, in the
// order the program builder gave them.
__device__ inline uint32_t round_key(int s, int l_node, int r_node, int cat) {
    return mix32((uint32_t)s * 0x9e3779b1u ^ ((uint32_t)l_node * 0x85ebca77u + (uint32_t)r_node * 0x27d4eb2fu +
                                              (uint32_t)cat * 0xc2b2ae3du));
}
__device__ inline uint32_t round_key(int s, const Op& op, int cat) {
    return (op.flags & OP_SWAP) ? round_key(s, op.r_node, op.l_node, cat) : round_key(s, op.l_node, op.r_node, cat);
}

__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Compensated sum of partial[0, n) by one wave in a fixed order: lane l adds
// partial[l], partial[l + 64], ... with Kahan compensation, then the lanes'
// (sum, error) pairs meet in a two-sum butterfly.  Every lane returns the
// same total (two-sum gives the exact error whichever operand comes first).
// Adds b + bl to the compensated pair (hi, lo) by Knuth's two-sum.
__device__ inline void two_sum_add(double& hi, double& lo, double b, double bl) {
    const double t = hi + b;
    const double bb = t - hi;
    const double err = (hi - (t - bb)) + (b - bb);
    hi = t;
    lo = (lo + bl) + err;
}

// Threads of the compensated sum: one workgroup per program.
constexpr int kSumThreads = 256;

// Compensated fixed-order sum of partial[0, n) by one workgroup of
// kSumThreads.  Each thread issues its loads eight at a time before adding
// them (Kahan), so their latencies overlap instead of queueing; the threads'
// (sum, error) pairs then meet by two-sum, in a butterfly within each wave
// and in LDS across the waves.  Thread 0 returns the total.
__device__ inline double block_total(const double* partial, int n) {
    double s = 0.0, comp = 0.0;
    for (int base = threadIdx.x; base < n; base += 8 * kSumThreads) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int i = base + u * kSumThreads;
            v[u] = i < n ? partial[i] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            if (base + u * kSumThreads < n) {
                const double x = v[u] - comp;
                const double t = s + x;
                comp = (t - s) - x;
                s = t;
            }
        }
    }
    double hi = s, lo = -comp;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) two_sum_add(hi, lo, __shfl_xor(hi, o, 64), __shfl_xor(lo, o, 64));
    __shared__ double red_hi[kSumThreads / 64], red_lo[kSumThreads / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) { red_hi[wid] = hi; red_lo[wid] = lo; }
    __syncthreads();
    hi = red_hi[0];
    lo = red_lo[0];
#pragma unroll
    for (int w = 1; w < kSumThreads / 64; ++w) two_sum_add(hi, lo, red_hi[w], red_lo[w]);
    return hi + lo;
}

// Sums v over the workgroup (any multiple of 64 threads); thread 0 stores it.
__device__ inline void block_sum_store(double v, double* dst) {
    __shared__ double red[16];
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) red[wid] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red[w];
        *dst = t;
    }
}

// The top launch of a tail split (its workgroups are the last blocks to
// finish an evaluation) folds in the final sum: each workgroup, once its
// partial is stored, counts itself in; the last one sums every block's
// partial in block_total's order (so the total equals final_sum_kernel's),
// writes it to the host-mapped total and returns the counter to zero.  The
// fences publish the few partials this launch wrote and drop stale copies
// before the sum reads them.  Workgroups of kSumThreads threads only.
__device__ inline void fold_total(const double* partial, int n_part, unsigned* counter, double* total) {
    __shared__ int last;
    if (threadIdx.x == 0) {
        __threadfence();
        last = atomicAdd(counter, 1u) == gridDim.x - 1;
        if (last) __threadfence();
    }
    __syncthreads();
    if (last) {
        const double t = block_total(partial, n_part);
        if (threadIdx.x == 0) {
            *total = t;
            *counter = 0u;
        }
    }
}

// Bit j of a tip's state set as 1.0 or 0.0: a bit-field extract and an
// integer-to-double conversion.  (Building the double's high word from a
// mask needs a third instruction to copy the zero low word into each
// register pair; the conversion form ran the C2 kernel 1 % faster.)
template <int J>
__device__ inline double tip_bit(unsigned code) {
    return (double)((code >> J) & 1u);
}

// frexp exponent of the largest of four values (0 if none is positive): as
// signed integers the high words of positive doubles order like the values,
// and those of negative ones (P(t) from the eigen system can leave tiny
// negative partials) are negative, so they never win.
__device__ inline int max_exponent(const double (&L)[4]) {
    const int h = max(max(__double2hiint(L[0]), __double2hiint(L[1])), max(__double2hiint(L[2]), __double2hiint(L[3])));
    int e = (h >> 20) - 1022;
    if (h < 0x00100000) {   // no positive normal value: zero, subnormal or negative maximum
        const double mx = fmax(fmax(L[0], L[1]), fmax(L[2], L[3]));
        e = 0;
        if (mx > 0.0) (void)frexp(mx, &e);
    }
    return e;
}

// The largest high word of four doubles, as signed integers: for positive
// values it orders like the values (negative ones never win).
__device__ inline int max_hi_word(const double (&L)[4]) {
    return max(max(__double2hiint(L[0]), __double2hiint(L[1])), max(__double2hiint(L[2]), __double2hiint(L[3])));
}

// High word of 2^-256: carried DNA partials whose largest entry is below it
// are rescaled (walk4).
constexpr int kLazyRescaleHi = (1023 - 256) << 20;

// One child's operand for a walker op, by source kind (compile-time).
// With TL the tip's state set comes from the workgroup's LDS copy, one byte
// per site holding 16 x code, which is the byte offset of the code's entry
// in tipv: states 0, 1 at tipv[code], states 2, 3 at tipv[16 + code] (two
// LDS reads addressed by the byte itself, where the bit-by-bit decode takes
// four bit-field extracts and four conversions).
template <int K, typename T, int C, bool TL>
__device__ inline void child4(int idx, const uint8_t* __restrict__ tips, const uint8_t* tl, const double2* tipv,
                              const T* plane, size_t slot_stride, const int* scl, int m_pad, int s, int lane, int cat,
                              const double (&cA)[4], int sA, double2 (&xb)[C][2][64], int sB, double (&x)[4], int& sc) {
    if (K == SRC_TIP) {
        if (TL) {
            const unsigned off = tl[idx * 64 + lane];
            const double2 a = *reinterpret_cast<const double2*>(reinterpret_cast<const char*>(tipv) + off);
            const double2 b = *reinterpret_cast<const double2*>(reinterpret_cast<const char*>(tipv + 16) + off);
            x[0] = a.x; x[1] = a.y; x[2] = b.x; x[3] = b.y;
        } else {
            const unsigned code = tips[(size_t)idx * m_pad + s];
            x[0] = tip_bit<0>(code);
            x[1] = tip_bit<1>(code);
            x[2] = tip_bit<2>(code);
            x[3] = tip_bit<3>(code);
        }
    } else if (K == SRC_CLV) {
        load4(plane + idx * slot_stride, x);
        sc += scl[((size_t)idx * C + cat) * m_pad + s];
    } else if (K == SRC_CARRY_A) {
#pragma unroll
        for (int j = 0; j < 4; ++j) x[j] = cA[j];
        sc += sA;
    } else {
        const double2 a = xb[cat][0][lane], b = xb[cat][1][lane];
        x[0] = a.x; x[1] = a.y; x[2] = b.x; x[3] = b.y;
        sc += sB;
    }
}

// L = (P_l x_l) * (P_r x_r) for one op whose child kinds are known at compile
// time.  The walker switches once per op over the (left, right) kinds into
// one straight-line body each; testing the kinds one by one costs the
// compiler's flag-and-branch glue on every op (about 10 % of the C2 kernel).
template <int LK, int RK, typename T, int C, bool TL>
__device__ inline void children4(const Op& op, const uint8_t* __restrict__ tips, const uint8_t* tl,
                                 const double2* tipv, const T* plane,
                                 size_t slot_stride, const int* scl, int m_pad, int s, int lane, int cat,
                                 const double (&cA)[4], int sA, double2 (&xb)[C][2][64], int sB,
                                 const double (&Pl)[16], const double (&Pr)[16], double (&L)[4], int& sc) {
    double x[4];
    child4<LK, T, C, TL>(op.l_idx, tips, tl, tipv, plane, slot_stride, scl, m_pad, s, lane, cat, cA, sA, xb, sB, x,
                         sc);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        double a = 0.0;
#pragma unroll
        for (int j = 0; j < 4; ++j) a = fma(Pl[i * 4 + j], x[j], a);
        L[i] = a;
    }
    if (RK != SRC_NONE) {
        child4<RK, T, C, TL>(op.r_idx, tips, tl, tipv, plane, slot_stride, scl, m_pad, s, lane, cat, cA, sA, xb, sB,
                             x, sc);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            double a = 0.0;
#pragma unroll
            for (int j = 0; j < 4; ++j) a = fma(Pr[i * 4 + j], x[j], a);
            L[i] *= a;
        }
    }
}

// The pruning walker for S <= 4 (DNA, binary; padded states carry zero rows/cols of P).
//
// A workgroup owns 64 consecutive sites; wave c of the workgroup computes rate
// category c for them, so the P matrices a wave needs are uniform across it
// (scalar loads) and each lane holds one float4 partial.  Each (node, site,
// category) partial carries its own power-of-two scaler, so the waves of a
// workgroup walk the program without waiting for each other; they meet only
// at the root, where the categories' sums are brought to a common exponent.
// A stored partial is scaled so its largest entry is in [0.5, 1) (the
// scaler is the frexp exponent of its maximum); a carried one is rescaled
// only when it drops below 2^-256.
// walk4 runs ops[k0, k1) for site s and category cat; when the last op is
// the root (a program's root op is always its last) it sets
// root_lc = w_cat sum_i rw_i L_i and root_sc, the category's scaler, so that
// its share of the site likelihood is root_lc 2^root_sc (the root is read
// from carry A after the loop, so the loop carries no root values).
// With TL the tip codes of the workgroup's 64 sites come from LDS (tl: site
// j of taxon t at tl[64 t + j], as 16 x code) and their 0/1 state values from
// the LDS table tipv (child4).
// Both children's P matrices of an op (64 SGPRs; their scalar loads
// overlap), by 32-bit byte offsets from this category's first matrix (three
// scalar instructions per address instead of five).
template <int C>
__device__ __forceinline__ void load_p2(const Op& op, const char* Pcat, double (&Pl)[16], double (&Pr)[16]) {
    const double* pl = reinterpret_cast<const double*>(Pcat + (uint32_t)op.l_node * (uint32_t)(C * 128));
    const double* pr = reinterpret_cast<const double*>(Pcat + (uint32_t)op.r_node * (uint32_t)(C * 128));
#pragma unroll
    for (int i = 0; i < 16; ++i) { Pl[i] = pl[i]; Pr[i] = pr[i]; }
}

// One op of walk4 (ops[k] for site s and category cat).  Carry A comes in
// as cin / sin and the result leaves as cout / sout: walk4 alternates two
// register sets, so the result is written where the next op reads it, with
// no copy.  With PIPE the op's P arrive in Pl / Pr, and the next op's (nx)
// are loaded into nPl / nPr once this op's products are done: the P
// registers are free by then, and the loads run under the rest of the op
// instead of at the start of the next.
template <typename T, int C, bool TL, bool PIPE>
__device__ __forceinline__ void walk4_op(const Op op, const uint8_t* __restrict__ tips, const uint8_t* tl,
                                         const double2* tipv, const T* plane, size_t slot_stride, int* scl,
                                         const char* Pcat, int m_pad, int s, int lane, int cat,
                                         double2 (&xb)[C][2][64], const double (&cin)[4], int sin, int& sB,
                                         double (&cout)[4], int& sout, const double (&Pin_l)[16],
                                         const double (&Pin_r)[16], const Op& nx, double (&nPl)[16],
                                         double (&nPr)[16]) {
    double Pl[16], Pr[16];
    if constexpr (PIPE) {
#pragma unroll
        for (int i = 0; i < 16; ++i) { Pl[i] = Pin_l[i]; Pr[i] = Pin_r[i]; }
    } else {
        load_p2<C>(op, Pcat, Pl, Pr);
    }
    // child inputs: one straight-line body per (left kind, right kind); the
    // commonest shapes by their one-hot hint bits first (single-bit tests,
    // which the compiler keeps apart from the switch's compare tree)
    double L[4];
    int sc = 0;
    if (op.flags & OP_SH_AT)
        children4<SRC_CARRY_A, SRC_TIP, T, C, TL>(op, tips, tl, tipv, plane, slot_stride, scl, m_pad, s, lane, cat,
                                                  cin, sin, xb, sB, Pl, Pr, L, sc);
    else if (op.flags & OP_SH_TT)
        children4<SRC_TIP, SRC_TIP, T, C, TL>(op, tips, tl, tipv, plane, slot_stride, scl, m_pad, s, lane, cat, cin,
                                              sin, xb, sB, Pl, Pr, L, sc);
    else if (op.flags & OP_SH_AB)
        children4<SRC_CARRY_A, SRC_CARRY_B, T, C, TL>(op, tips, tl, tipv, plane, slot_stride, scl, m_pad, s, lane, cat,
                                                      cin, sin, xb, sB, Pl, Pr, L, sc);
    else if (op.flags & OP_SH_AC)
        children4<SRC_CARRY_A, SRC_CLV, T, C, TL>(op, tips, tl, tipv, plane, slot_stride, scl, m_pad, s, lane, cat,
                                                  cin, sin, xb, sB, Pl, Pr, L, sc);
    else switch (op.l_kind * 5 + op.r_kind) {
#define PHYLO_SHAPE(A, B)                                                                                   \
    case A * 5 + B:                                                                                       \
        children4<A, B, T, C, TL>(op, tips, tl, tipv, plane, slot_stride, scl, m_pad, s, lane, cat, cin, sin, xb, \
                                  sB, Pl, Pr, L, sc);                                                     \
        break;
        // the other shapes of children in kind order (carry A, tip, carry B,
        // stored, none)
        PHYLO_SHAPE(2, 4) PHYLO_SHAPE(0, 3) PHYLO_SHAPE(0, 1) PHYLO_SHAPE(0, 4)
        PHYLO_SHAPE(3, 1) PHYLO_SHAPE(3, 4) PHYLO_SHAPE(1, 1)
        default: PHYLO_SHAPE(1, 4)      // (no other shape occurs: one carry A and one carry B at a time)
#undef PHYLO_SHAPE
    }
    if constexpr (PIPE) load_p2<C>(nx, Pcat, nPl, nPr);
    // Carried results are FP64 and rescaled lazily: only when the
    // largest entry falls below 2^-256 (rare; a lane-divergent branch
    // costs the others one compare), so an op's children are each at
    // least 2^-256 and their product stays far above the FP64 normal
    // range.  (Rescaling by a power of two is exact, so where it
    // happens changes no value.)
    if (max_hi_word(L) < kLazyRescaleHi) {
        const int e = max_exponent(L);
        sc += e;
#pragma unroll
        for (int i = 0; i < 4; ++i) L[i] = ldexp(L[i], -e);
    }
    if (op.out_slot >= 0) {
        // a stored partial is brought to [0.5, 1) with its own scaler
        // (the carried value keeps its scale, so programs that store
        // and programs that carry a node compute the same values), then
        // rounded to T: most results are only carried (the program
        // stores what it reads back), so the rounding sits here, on the
        // few stored ones
        const int e = max_exponent(L);
        const uint32_t key = round_key(s, op, cat);
        T r[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) r[i] = (T)round_cut(ldexp(L[i], -e), key + 0x9e3779b9u * i, T());
        store4(const_cast<T*>(plane) + op.out_slot * slot_stride, r);
        scl[((size_t)op.out_slot * C + cat) * m_pad + s] = sc + e;
    }
    // carry A takes every result (after an op to carry B nothing reads
    // the old A), so no select between the carries
#pragma unroll
    for (int i = 0; i < 4; ++i) cout[i] = L[i];
    sout = sc;
    const bool to_b = op.flags & OP_TO_B;
    if (to_b) {
        xb[cat][0][lane] = make_double2(L[0], L[1]);
        xb[cat][1][lane] = make_double2(L[2], L[3]);
    }
    sB = to_b ? sc : sB;
}

// PAIRS: two ops per loop turn (the whole-tree and NNI launches); without
// it one op per turn, the result copied into carry A (the batch kernel,
// where the second copy of the op body would cost it a resident wave).  The
// values are the same either way.
template <typename T, int C, bool TL, bool PAIRS = true>
__device__ inline void walk4(const Op* __restrict__ ops, int k0, int k1, const uint8_t* __restrict__ tips,
                             const uint8_t* tl, const double2* tipv, T* clv, int* scl, const double* __restrict__ P,
                             const double* __restrict__ roots, int m_pad, int s, int lane, int cat,
                             double2 (&xb)[C][2][64], double& root_lc, int& root_sc) {
    // carried results stay in FP64: only stored partials are rounded to T
    // (carry B, the one result waiting while a sibling subtree is walked,
    // sits in LDS: xb[cat][][lane], written and read by the same lane)
    double cA[4] = {0.0, 0.0, 0.0, 0.0};
    int sA = 0, sB = 0;
    const T* plane = clv + ((size_t)cat * m_pad + s) * 4;          // + slot * C * m_pad * 4
    const size_t slot_stride = (size_t)C * m_pad * 4;
    const char* Pcat = reinterpret_cast<const char*>(P + (size_t)cat * 16);
    double cB2[4] = {0.0, 0.0, 0.0, 0.0};
    int sA2 = 0;
    int k = k0;
    if constexpr (!PAIRS) {
        double unused[16];
        for (; k < k1; ++k) {
            walk4_op<T, C, TL, false>(ops[k], tips, tl, tipv, plane, slot_stride, scl, Pcat, m_pad, s, lane, cat, xb,
                                      cA, sA, sB, cB2, sA2, unused, unused, ops[k], unused, unused);
#pragma unroll
            for (int i = 0; i < 4; ++i) cA[i] = cB2[i];
            sA = sA2;
        }
    } else {
        // two ops per turn, carry A alternating between cA and cB2, and the
        // P matrices between PA and PB.  Each op brings in the next op
        // (its fields loaded one op ahead, its P once the products are
        // done), so the next op starts without a scalar-load wait.  The op
        // buffer holds a zero op after the last program (node 0's P are
        // loaded for it, never used), so ops[k1] is always readable.
        double PA_l[16], PA_r[16], PB_l[16], PB_r[16];
        const Op* op = ops + k0;
        const Op* const op_end = ops + k1;
        Op cur = op[0];
        load_p2<C>(cur, Pcat, PA_l, PA_r);
        for (; op + 1 < op_end; op += 2) {
            const Op nx = op[1];
            walk4_op<T, C, TL, true>(cur, tips, tl, tipv, plane, slot_stride, scl, Pcat, m_pad, s, lane, cat, xb,
                                     cA, sA, sB, cB2, sA2, PA_l, PA_r, nx, PB_l, PB_r);
            cur = op[2];
            walk4_op<T, C, TL, true>(nx, tips, tl, tipv, plane, slot_stride, scl, Pcat, m_pad, s, lane, cat,
                                     xb, cB2, sA2, sB, cA, sA, PB_l, PB_r, cur, PA_l, PA_r);
        }
        if (op < op_end) {  // an odd op count: the last op, its result back in cA
            walk4_op<T, C, TL, true>(cur, tips, tl, tipv, plane, slot_stride, scl, Pcat, m_pad, s, lane, cat, xb,
                                     cA, sA, sB, cB2, sA2, PA_l, PA_r, cur, PB_l, PB_r);
#pragma unroll
            for (int i = 0; i < 4; ++i) cA[i] = cB2[i];
            sA = sA2;
        }
    }
    root_lc = 0.0;
    root_sc = 0;
    if (k1 > k0 && (ops[k1 - 1].flags & OP_ROOT)) {
        // the root partial (carry A) is brought to [0.5, 1) first, so the
        // site value does not depend on where the walk rescaled (a program
        // that reads a stored child and one that carries it reach the root
        // with the same mantissas)
        const int e = max_exponent(cA);
        double lc = 0.0;
#pragma unroll
        for (int i = 0; i < 4; ++i) lc = fma(roots[i], ldexp(cA[i], -e), lc);
        root_lc = lc * roots[kRootCatOff + cat];
        root_sc = sA + e;
    }
}

// Site lnL from the categories' root sums (lc, sc) held by the C waves of a
// workgroup (wave = category): the sums meet in LDS at their largest
// exponent.  Every wave returns the same value.  REUSE: the caller writes
// xroot / xsc again later, so the waves meet once more after reading.
template <int C, bool REUSE = true>
__device__ inline double combine_categories(double lc, int sc, int cat, int lane, double (&xroot)[C][64],
                                            int (&xsc)[C][64]) {
    if (C > 1) {
        xroot[cat][lane] = lc;
        xsc[cat][lane] = sc;
        __syncthreads();
        int M = xsc[0][lane];
#pragma unroll
        for (int c = 1; c < C; ++c) M = max(M, xsc[c][lane]);
        lc = 0.0;
#pragma unroll
        for (int c = 0; c < C; ++c) lc += ldexp(xroot[c][lane], xsc[c][lane] - M);
        if (REUSE) __syncthreads();
        sc = M;
    }
    return log(lc) + sc * kLn2;
}

// Taxa whose tip codes a DNA workgroup stages in LDS (one byte per site,
// sized at launch).
constexpr int kTipLdsTaxa = 256;
inline size_t tip_lds_bytes(int n_taxa) { return (size_t)n_taxa * 64; }

// Plain launch (prog_main < 0): grid = (m_pad / 64, n_programs); workgroup
// (x, y) runs ops[prog_begin[y] .. prog_begin[y+1]) over site block x.
// Program-major launch (prog_main = -2, n_main site blocks, n_tail
// programs, blk0 programs a workgroup; 1-D grid of 8 ceil(n_main / 8)
// ceil(n_tail / blk0) workgroups): workgroup i runs on XCD i % 8
// (workgroups are dealt to the XCDs in turn), and XCD k's workgroups take
// site blocks k, k + 8, ... each with every group of blk0 consecutive
// programs in turn, so the programs of one site block read its stored
// partials through one L2.  Program p over block b writes
// partial[p n_main + b], as a plain launch does.
// Split launch (prog_main >= 0, grid.y = 1; build_split_programs): workgroups
// x < n_main run program prog_main over site block blk0 + x; the others run
// piece program prog_piece + (x - n_main) / n_tail over site block
// tail0 + (x - n_main) % n_tail, which has no root: they store their results
// and leave.  The split's second launch (the tail's top program) passes
// fold_counter: its last workgroup also writes the evaluation's total
// (fold_total) when the workgroup has kSumThreads threads.
// TL (n_taxa <= kTipLdsTaxa): the workgroup first copies its 64 sites of
// every tip into LDS, so no op of the walk waits on a global tip load.
template <typename T, int C, bool TL>
__global__ __launch_bounds__(64 * C) void prune4_kernel(
    const Op* __restrict__ ops, const int* __restrict__ prog_begin, const uint8_t* __restrict__ tips,
    T* clv, int* scl, const double* __restrict__ P, const double* __restrict__ roots,
    const double* __restrict__ site_w, double* __restrict__ site_lnl, double* __restrict__ partial,
    int m, int m_pad, int n_taxa, int n_main, int prog_main, int blk0, int prog_piece, int tail0, int n_tail,
    unsigned* fold_counter, double* fold_out) {
    __shared__ int xsc[C][64];
    __shared__ double xroot[C][64];
    __shared__ double2 xb[C][2][64];
    extern __shared__ __align__(16) uint8_t tl[];   // TL: n_taxa x 64 bytes (launch size)
    __shared__ double2 tipv[TL ? 32 : 1];
    const int lane = threadIdx.x & 63;
    const int cat = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int y = blockIdx.y;
    int bx = blockIdx.x, prog = y;
    bool has_root = true;
    if (prog_main >= 0) {
        if ((int)blockIdx.x < n_main) {
            bx = blk0 + (int)blockIdx.x;
            prog = prog_main;
        } else {
            const int r = (int)blockIdx.x - n_main;
            bx = tail0 + r % n_tail;
            prog = prog_piece + r / n_tail;
            has_root = false;
        }
    } else if (prog_main == -2) {
        const int slot = (int)blockIdx.x >> 3, n_pg = (n_tail + blk0 - 1) / blk0;
        prog = slot % n_pg * blk0;
        bx = slot / n_pg * 8 + ((int)blockIdx.x & 7);
        if (bx >= n_main) return;        // the last group's padding (uniform per workgroup)
    }
    const int s = bx * 64 + lane;
    if (TL) {
        if (threadIdx.x < 32) {          // 0/1 values of states 2h, 2h + 1 of each code
            const unsigned code = threadIdx.x & 15, h = (threadIdx.x >> 4) * 2;
            tipv[threadIdx.x] = make_double2((double)((code >> h) & 1u), (double)((code >> (h + 1)) & 1u));
        }
        const uint8_t* blk = tips + (size_t)bx * 64;
        for (int i = threadIdx.x; i < n_taxa * 16; i += 64 * C) {      // four sites a thread
            const uint32_t w = *reinterpret_cast<const uint32_t*>(blk + (size_t)(i >> 4) * m_pad + 4 * (i & 15));
            reinterpret_cast<uint32_t*>(tl)[i] = (w << 4) & 0xF0F0F0F0u;
        }
        __syncthreads();
    }
    if constexpr (!TL) {
        if (prog_main == -2) {
            // blk0 programs of one site block, one after another
            const int p_end = min(prog + blk0, n_tail);
            for (int p = prog; p < p_end; ++p) {
                double lc;
                int sc;
                walk4<T, C, false>(ops, prog_begin[p], prog_begin[p + 1], tips, tl, tipv, clv, scl, P, roots, m_pad,
                                   s, lane, cat, xb, lc, sc);
                const double lnl = combine_categories<C, true>(lc, sc, cat, lane, xroot, xsc);
                const bool live = s < m && cat == 0;
                if (live && site_lnl) site_lnl[(size_t)p * m + s] = lnl;
                if (cat == 0) {
                    const double v = wave_sum(live ? (site_w ? site_w[s] : 1.0) * lnl : 0.0);
                    if (lane == 0) partial[(size_t)p * n_main + bx] = v;
                }
            }
            return;
        }
    }
    double lc;
    int sc;
    walk4<T, C, TL>(ops, prog_begin[prog], prog_begin[prog + 1], tips, tl, tipv, clv, scl, P, roots, m_pad, s, lane,
                    cat, xb, lc, sc);
    if (!has_root) return;    // a piece: its stores are its result (uniform per workgroup)
    const double lnl = combine_categories<C, false>(lc, sc, cat, lane, xroot, xsc);
    const bool live = s < m && cat == 0;
    if (live && site_lnl) site_lnl[(size_t)y * m + s] = lnl;
    const double w = live ? (site_w ? site_w[s] : 1.0) : 0.0;
    double* dst = partial + (prog_main >= 0 ? (size_t)bx : (size_t)y * gridDim.x + blockIdx.x);
    if (64 * C == kSumThreads && fold_counter) {
        block_sum_store(live ? w * lnl : 0.0, dst);
        fold_total(partial, blk0 + n_main, fold_counter, fold_out);
    } else if (cat == 0) {
        // only the category-0 wave holds site values: its sum is the
        // block's (the other waves' zeros would add nothing)
        const double v = wave_sum(live ? w * lnl : 0.0);
        if (lane == 0) *dst = v;
    }
}

// Batched scoring: many independent problems (and their NNI programs) in one
// persistent launch.  Workgroup g takes work items (instance, program, block
// of 64 sites) g, g + grid, g + 2*grid, ... of a list sorted by decreasing
// cost; the loop counter is uniform, so every wave runs the same barriers and
// every workgroup reaches the exit.  Each item stores its partial sum at
// items[].out.  (A thread-0 atomic work queue is deliberately not used: the
// compiler merges a thread-0 queue fetch at the top of the loop with the
// thread-0 partial store at its tail into one divergent region, and the other
// lanes then loop on a stale item.)
struct BatchInst {
    long tips_off;    // into the uint8 tip-code buffer
    long clv_off;     // elements into the partial buffer
    long scl_off;     // into the scaler buffer
    long P_off;       // doubles into the P buffer
    int roots_off;    // doubles into the root-weight buffer
    int m, m_pad, pad;
};
struct WorkItem {
    int inst, prog, block, out;
};

// Category tiling CW: CW = C puts the C categories on the C waves of a
// workgroup (they meet in LDS at the root); CW = 1 has one wave walk the
// program once per category and combine the root sums in registers, with no
// barrier at all (a better fit for short programs such as NNI variants).
template <typename T, int C, int CW>
__global__ __launch_bounds__(64 * CW) void batch4_kernel(const WorkItem* __restrict__ items, int n_items,
                                                         const BatchInst* __restrict__ inst,
                                                         const int* __restrict__ prog_begin,
                                                         const Op* __restrict__ ops, const uint8_t* __restrict__ tips,
                                                         T* clv, int* scl, const double* __restrict__ P,
                                                         const double* __restrict__ roots,
                                                         double* __restrict__ partial, int blocks_per_item) {
    static_assert(CW == C || CW == 1, "categories per workgroup: all on separate waves, or one wave");
    __shared__ int xsc[C][64];
    __shared__ double xroot[C][64];
    __shared__ double2 xb[C][2][64];
    const int lane = threadIdx.x & 63;
    const int wcat = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (int it = blockIdx.x; it < n_items; it += gridDim.x) {
        const WorkItem w = items[it];
        const BatchInst I = inst[w.inst];
        const int k0 = prog_begin[w.prog], k1 = prog_begin[w.prog + 1];
        double acc = 0.0;
        for (int q = 0; q < blocks_per_item; ++q) {
            const int s = (w.block + q) * 64 + lane;
            if ((w.block + q) * 64 >= I.m_pad) break;
            double lnl;
            if (CW == C) {
                double lc;
                int sc;
                walk4<T, C, false, false>(ops, k0, k1, tips + I.tips_off, nullptr, nullptr, clv + I.clv_off,
                                          scl + I.scl_off, P + I.P_off, roots + I.roots_off, I.m_pad, s, lane, wcat,
                                          xb, lc, sc);
                lnl = combine_categories<C>(lc, sc, wcat, lane, xroot, xsc);
            } else {
                double lc[C];
                int sc[C];
                for (int cat = 0; cat < C; ++cat)
                    walk4<T, C, false, false>(ops, k0, k1, tips + I.tips_off, nullptr, nullptr,
                                              clv + I.clv_off, scl + I.scl_off, P + I.P_off, roots + I.roots_off,
                                              I.m_pad, s, lane, cat, xb, lc[cat], sc[cat]);
                int M = sc[0];
#pragma unroll
                for (int c = 1; c < C; ++c) M = max(M, sc[c]);
                double t = 0.0;
#pragma unroll
                for (int c = 0; c < C; ++c) t += ldexp(lc[c], sc[c] - M);
                lnl = log(t) + M * kLn2;
            }
            if (s < I.m && wcat == 0) acc += lnl;
        }
        block_sum_store(acc, partial + w.out);
    }
}

// Segmented compensated sums: out[k] = sum partial[seg[k] .. seg[k+1]).
__global__ __launch_bounds__(64) void segment_sum_kernel(const double* __restrict__ partial,
                                                        const int* __restrict__ seg, int n_seg,
                                                        double* __restrict__ out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_seg) return;
    double s = 0.0, comp = 0.0;
    for (int i = seg[k]; i < seg[k + 1]; ++i) {
        const double x = partial[i] - comp;
        const double t = s + x;
        comp = (t - s) - x;
        s = t;
    }
    out[k] = s;
}

// ---------------------------------------------------------------------------
// Wide-state walker on the matrix cores (4 < S <= 32, FP64 arithmetic).
//
// One wave = G groups of 16 sites x one rate category; a workgroup holds the
// C category waves of the same G site groups.  Each child product is a GEMM
//     out[state row][site] = sum_k P[row][k] * L[k][site]
// done as v_mfma_f64_16x16x4f64 steps (A = P: 16 rows x 4 k, B = L: 4 k x 16
// sites, D = 16 rows x 16 sites; D[i][j] sits in lane j + 16 (i % 4),
// register i / 4).  Lane l serves site l % 16 and lane group g = l / 16; slot
// t of a lane holds state 4t + g, and row i of tile T is state 16T + i, so an
// MFMA result (tile T, register r) is slot 4T + r of the next op's operand
// with no shuffles.  V = ceil(S / 4) slots cover the S states; tile 1 exists
// for S > 16 (for 17..20 states as v_mfma_f64_4x4x4f64 blocks, see
// pmfma_kernel).  Partials are stored as [slot][cat][site group][V][64 lanes]
// (every load and store a full row), P as [node][cat][tile][V][64] doubles
// (pmfma_kernel).  FP64 matrix cores rather than FP32 ones: the FP32 MFMA
// dot products truncate small addends (about -0.1 ulp per 5-step chain on
// transition-matrix-shaped operands), which biases every site the same way.
// Scalers as in the DNA walker: one int32 power of two per (node, category,
// site), the categories meeting only at the root.
// ---------------------------------------------------------------------------

typedef double mfma_d4 __attribute__((ext_vector_type(4)));

__host__ __device__ constexpr int mfma_slots(int S) { return (S + 3) / 4; }

// Pm[node][cat][T][t][lane] = P[16T + lane % 16][4t + lane / 16] (0 outside S),
// the A operand of v_mfma_f64_16x16x4f64 for row tile T and K-step t.  With
// V = 5 (17..20 states) tile 1 holds only rows 16..19 and is instead the A
// operand of v_mfma_f64_4x4x4f64 (four 4x4 blocks, one per 4 sites): lane l
// = 4b + 16k + i holds P[16 + i][4t + k] for every block b.  That product
// lands as D[lane j + 16 i] = row 16 + i of site j, the operand slot of
// states 16..19, with no padded rows.
__global__ void pmfma_kernel(const double* __restrict__ P, int S, int Sp, int C, int n_nodes, int V, int TN,
                             double* __restrict__ Pm) {
    const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long total = (long)n_nodes * C * TN * V * 64;
    if (gid >= total) return;
    const int l = gid % 64;
    const int t = (gid / 64) % V;
    const int T = (gid / (64L * V)) % TN;
    const long mat = gid / (64L * V * TN);          // node * C + cat
    const bool short_tile = V == 5 && T == 1;
    const int row = 16 * T + (short_tile ? l % 4 : l % 16), col = 4 * t + l / 16;
    Pm[gid] = (row < S && col < S) ? P[(mat * Sp + row) * Sp + col] : 0.0;
}

// PTm[node][cat][k][i] = P[i][k] for k < S, PTm[..][S][i] = sum_q P[i][q]
// (the product with an all-states tip), i < 4V (0 for i >= S): the columns a
// single-state or all-states tip selects.
__global__ void ptm_kernel(const double* __restrict__ P, int S, int Sp, int C, int n_nodes, int W,
                           double* __restrict__ PTm) {
    const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long total = (long)n_nodes * C * (S + 1) * W;
    if (gid >= total) return;
    const int i = gid % W;
    const int k = (gid / W) % (S + 1);
    const long mat = gid / ((long)W * (S + 1));
    const double* row = P + (mat * Sp + i) * Sp;
    double v = 0.0;
    if (i < S) {
        if (k < S) v = row[k];
        else
            for (int q = 0; q < S; ++q) v += row[q];
    }
    PTm[gid] = v;
}

__device__ inline double to_d(float x) { return (double)x; }
__device__ inline double to_d(double x) { return x; }

// Largest state value of a site in the matrix walker's layout (states 4t + g
// in slot t of lane group g, so the maximum runs over the slots and then the
// four lane groups: lanes l ^ 16, l ^ 32), and its frexp exponent (0 when no
// state is positive).  Every lane of the wave must take part.
template <int V>
__device__ inline double site_max(const double (&v)[V]) {
    double mx = 0.0;
#pragma unroll
    for (int t = 0; t < V; ++t) mx = fmax(mx, v[t]);
    mx = fmax(mx, __shfl_xor(mx, 16, 64));
    return fmax(mx, __shfl_xor(mx, 32, 64));
}
template <int V>
__device__ inline int site_exponent(const double (&v)[V]) {
    const double mx = site_max<V>(v);
    int e = 0;
    if (mx > 0.0) (void)frexp(mx, &e);
    return e;
}

// One child for the G site groups of a wave: acc[q][T] = P . x for tile T,
// x in operand order.  The P operand of each K-step is loaded once and used by
// every group.
template <typename T, int C, int V, int G, bool TL>
__device__ inline void mfma_child(int kind, int idx, int node, int cat, const int (&s)[G], int sg0, int n_sg,
                                  int lane, int g, int m_pad, int S, uint32_t full,
                                  const uint32_t* __restrict__ tips, const uint32_t* tl, const T* clv,
                                  size_t slot_stride,
                                  const int* scl, const double* __restrict__ Pm, const double* __restrict__ PTm,
                                  const double (&cA)[G][V], const int (&sA)[G], const double* cB,
                                  const int (&sB)[G], mfma_d4 (&acc)[G][2], int (&sc)[G]) {
    constexpr int TN = V > 4 ? 2 : 1;
    double x[G][V];
    if (kind == SRC_TIP) {
        uint32_t mk[G];
        bool simple = true;
#pragma unroll
        for (int q = 0; q < G; ++q) {
            mk[q] = TL ? tl[idx * 16 * G + 16 * q + (lane & 15)] : tips[(size_t)idx * m_pad + s[q]];
            simple = simple && (mk[q] == full || __popc(mk[q]) == 1);
        }
        // a wave whose sites are all single states or gaps selects columns of P
        if (__all(simple)) {
            const double* base = PTm + ((size_t)node * C + cat) * (S + 1) * (4 * V) + g;
#pragma unroll
            for (int q = 0; q < G; ++q) {
                const int col = mk[q] == full ? S : __ffs(mk[q]) - 1;
                const double* pt = base + (size_t)col * (4 * V);
#pragma unroll
                for (int t = 0; t < V; ++t) acc[q][t / 4][t % 4] = pt[4 * t];
            }
            return;
        }
#pragma unroll
        for (int q = 0; q < G; ++q)
#pragma unroll
            for (int t = 0; t < V; ++t) x[q][t] = (mk[q] >> (4 * t + g)) & 1u ? 1.0 : 0.0;
    } else if (kind == SRC_CLV) {
#pragma unroll
        for (int q = 0; q < G; ++q) {
            const T* src = clv + (size_t)idx * slot_stride + ((size_t)cat * n_sg + sg0 + q) * V * 64 + lane;
#pragma unroll
            for (int t = 0; t < V; ++t) x[q][t] = to_d(src[t * 64]);
            sc[q] += scl[((size_t)idx * C + cat) * m_pad + s[q]];
        }
    } else if (kind == SRC_CARRY_A) {
#pragma unroll
        for (int q = 0; q < G; ++q) {
#pragma unroll
            for (int t = 0; t < V; ++t) x[q][t] = to_d(cA[q][t]);
            sc[q] += sA[q];
        }
    } else {
#pragma unroll
        for (int q = 0; q < G; ++q) {
#pragma unroll
            for (int t = 0; t < V; ++t) x[q][t] = cB[(q * V + t) * 64];
            sc[q] += sB[q];
        }
    }
    const double* pm = Pm + ((size_t)node * C + cat) * TN * V * 64 + lane;
#pragma unroll
    for (int T2 = 0; T2 < TN; ++T2) {
#pragma unroll
        for (int q = 0; q < G; ++q) acc[q][T2] = mfma_d4{0.0, 0.0, 0.0, 0.0};
        if (V == 5 && T2 == 1) {
            // states 16..19: four 4x4x4 blocks give exactly the 4 rows
            double d[G];
#pragma unroll
            for (int q = 0; q < G; ++q) d[q] = 0.0;
#pragma unroll
            for (int t = 0; t < V; ++t) {
                const double a = pm[(V + t) * 64];
#pragma unroll
                for (int q = 0; q < G; ++q) d[q] = __builtin_amdgcn_mfma_f64_4x4x4f64(a, x[q][t], d[q], 0, 0, 0);
            }
#pragma unroll
            for (int q = 0; q < G; ++q) acc[q][1][0] = d[q];
            continue;
        }
#pragma unroll
        for (int t = 0; t < V; ++t) {
            const double a = pm[(T2 * V + t) * 64];
#pragma unroll
            for (int q = 0; q < G; ++q)
                acc[q][T2] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, x[q][t], acc[q][T2], 0, 0, 0);
        }
    }
}

// The P operand of one child for every row tile and K-step (Pm order: tile
// T, K-step t at register T V + t).
template <int C, int V>
__device__ __forceinline__ void load_pm(int node, int cat, int lane, const double* __restrict__ Pm,
                                        double (&pf)[(V > 4 ? 2 : 1) * V]) {
    constexpr int TN = V > 4 ? 2 : 1;
    const double* pm = Pm + ((size_t)node * C + cat) * TN * V * 64 + lane;
#pragma unroll
    for (int i = 0; i < TN * V; ++i) pf[i] = pm[i * 64];
}

// r = P . x for the G site groups with the P operand in registers: the same
// matrix instructions in the same order as mfma_child; r[q][t] is state
// 4t + g of group q's site.
template <int V, int G>
__device__ __forceinline__ void mfma_products(const double (&pf)[(V > 4 ? 2 : 1) * V], const double (&x)[G][V],
                                              double (&r)[G][V]) {
    constexpr int TN = V > 4 ? 2 : 1;
    mfma_d4 acc[G][2];
#pragma unroll
    for (int T2 = 0; T2 < TN; ++T2) {
        if (V == 5 && T2 == 1) {
            double d[G];
#pragma unroll
            for (int q = 0; q < G; ++q) d[q] = 0.0;
#pragma unroll
            for (int t = 0; t < V; ++t)
#pragma unroll
                for (int q = 0; q < G; ++q) d[q] = __builtin_amdgcn_mfma_f64_4x4x4f64(pf[V + t], x[q][t], d[q], 0, 0, 0);
#pragma unroll
            for (int q = 0; q < G; ++q) acc[q][1][0] = d[q];
            continue;
        }
#pragma unroll
        for (int q = 0; q < G; ++q) acc[q][T2] = mfma_d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int t = 0; t < V; ++t)
#pragma unroll
            for (int q = 0; q < G; ++q)
                acc[q][T2] = __builtin_amdgcn_mfma_f64_16x16x4f64(pf[T2 * V + t], x[q][t], acc[q][T2], 0, 0, 0);
    }
#pragma unroll
    for (int q = 0; q < G; ++q)
#pragma unroll
        for (int t = 0; t < V; ++t) r[q][t] = acc[q][t / 4][t % 4];
}

// A tip child's state masks for the G site groups, and whether every site of
// the wave is a single state or a gap (then the child's factor is a column
// of P).
template <int G, bool TL>
__device__ __forceinline__ bool tip_masks(int idx, const int (&s)[G], int lane, int m_pad, uint32_t full,
                                          const uint32_t* __restrict__ tips, const uint32_t* tl, uint32_t (&mk)[G]) {
    bool simple = true;
#pragma unroll
    for (int q = 0; q < G; ++q) {
        mk[q] = TL ? tl[idx * 16 * G + 16 * q + (lane & 15)] : tips[(size_t)idx * m_pad + s[q]];
        simple = simple && (mk[q] == full || __popc(mk[q]) == 1);
    }
    return __all(simple);
}

// One child's factor in a shape-specialised op: a tip's P columns (col) or
// products with its state indicators, a stored or carried partial's products.
template <int K, typename T, int C, int V, int G>
__device__ __forceinline__ void mfma_factor(bool col, const double (&cols)[G][V], const uint32_t (&mk)[G], int node,
                                            int cat, int lane, int g, const double* __restrict__ Pm,
                                            const double (&pf)[(V > 4 ? 2 : 1) * V], const T (&xr)[G][V],
                                            const double (&cA)[G][V], const double* cB, double (&r)[G][V]) {
    if constexpr (K == SRC_TIP) {
        if (col) {
#pragma unroll
            for (int q = 0; q < G; ++q)
#pragma unroll
                for (int t = 0; t < V; ++t) r[q][t] = cols[q][t];
            return;
        }
        double pt[(V > 4 ? 2 : 1) * V], x[G][V];
        load_pm<C, V>(node, cat, lane, Pm, pt);
#pragma unroll
        for (int q = 0; q < G; ++q)
#pragma unroll
            for (int t = 0; t < V; ++t) x[q][t] = (mk[q] >> (4 * t + g)) & 1u ? 1.0 : 0.0;
        mfma_products<V, G>(pt, x, r);
    } else if constexpr (K == SRC_CLV) {
        double x[G][V];
#pragma unroll
        for (int q = 0; q < G; ++q)
#pragma unroll
            for (int t = 0; t < V; ++t) x[q][t] = to_d(xr[q][t]);
        mfma_products<V, G>(pf, x, r);
    } else if constexpr (K == SRC_CARRY_A) {
        mfma_products<V, G>(pf, cA, r);
    } else {
        double x[G][V];
#pragma unroll
        for (int q = 0; q < G; ++q)
#pragma unroll
            for (int t = 0; t < V; ++t) x[q][t] = cB[(q * V + t) * 64];
        mfma_products<V, G>(pf, x, r);
    }
}

// An op of child shape (LK, RK) (kind order, RK != SRC_NONE): the tip masks
// are read first, then every load of the op (P operands, stored partials and
// their scalers, tip columns) is issued before the first product, so the two
// children's memory latencies overlap.  val = left factor x right factor and
// sc = the children's scalers, as the generic op computes them.
template <int LK, int RK, typename T, int C, int V, int G, bool TL>
__device__ __forceinline__ void mfma_shape(const Op& op, int cat, const int (&s)[G], int sg0, int n_sg, int lane,
                                           int g, int m_pad, int S, uint32_t full, const uint32_t* __restrict__ tips,
                                           const uint32_t* tl, const T* clv, size_t slot_stride, const int* scl,
                                           const double* __restrict__ Pm, const double* __restrict__ PTm,
                                           const double (&cA)[G][V], const int (&sA)[G], const double* cB,
                                           const int (&sB)[G], double (&val)[G][V], int (&sc)[G]) {
    constexpr int TNV = (V > 4 ? 2 : 1) * V;
    uint32_t mkL[G], mkR[G];
    bool colL = false, colR = false;
    if constexpr (LK == SRC_TIP) colL = tip_masks<G, TL>(op.l_idx, s, lane, m_pad, full, tips, tl, mkL);
    if constexpr (RK == SRC_TIP) colR = tip_masks<G, TL>(op.r_idx, s, lane, m_pad, full, tips, tl, mkR);
    double pL[TNV], pR[TNV];
    if constexpr (LK != SRC_TIP) load_pm<C, V>(op.l_node, cat, lane, Pm, pL);
    T xL[G][V], xR[G][V];
    const auto load_stored = [&](int idx, T (&xr)[G][V]) {
#pragma unroll
        for (int q = 0; q < G; ++q) {
            const T* src = clv + (size_t)idx * slot_stride + ((size_t)cat * n_sg + sg0 + q) * V * 64 + lane;
#pragma unroll
            for (int t = 0; t < V; ++t) xr[q][t] = src[t * 64];
            sc[q] += scl[((size_t)idx * C + cat) * m_pad + s[q]];
        }
    };
    if constexpr (LK == SRC_CLV) load_stored(op.l_idx, xL);
    if constexpr (RK == SRC_CLV) load_stored(op.r_idx, xR);
    double gL[G][V], gR[G][V];
    const auto tip_columns = [&](int node, const uint32_t (&mk)[G], double (&cols)[G][V]) {
        const double* base = PTm + ((size_t)node * C + cat) * (S + 1) * (4 * V) + g;
#pragma unroll
        for (int q = 0; q < G; ++q) {
            const int col = mk[q] == full ? S : __ffs(mk[q]) - 1;
            const double* pt = base + (size_t)col * (4 * V);
#pragma unroll
            for (int t = 0; t < V; ++t) cols[q][t] = pt[4 * t];
        }
    };
    if constexpr (LK == SRC_TIP)
        if (colL) tip_columns(op.l_node, mkL, gL);
    if constexpr (RK == SRC_TIP)
        if (colR) tip_columns(op.r_node, mkR, gR);
#pragma unroll
    for (int q = 0; q < G; ++q) {
        if constexpr (LK == SRC_CARRY_A || RK == SRC_CARRY_A) sc[q] += sA[q];
        if constexpr (LK == SRC_CARRY_B || RK == SRC_CARRY_B) sc[q] += sB[q];
    }
    mfma_factor<LK, T, C, V, G>(colL, gL, mkL, op.l_node, cat, lane, g, Pm, pL, xL, cA, cB, val);
    if constexpr (RK != SRC_TIP) load_pm<C, V>(op.r_node, cat, lane, Pm, pR);
    double r[G][V];
    mfma_factor<RK, T, C, V, G>(colR, gR, mkR, op.r_node, cat, lane, g, Pm, pR, xR, cA, cB, r);
#pragma unroll
    for (int q = 0; q < G; ++q)
#pragma unroll
        for (int t = 0; t < V; ++t) val[q][t] *= r[q][t];
}

// grid = (m_pad / (16 G), n_programs), block = 64 * C: wave c computes
// category c of the G 16-site groups starting at group G * x of site block
// x (blockIdx.x in a plain launch; the tail-split mapping of prune4_kernel
// otherwise).
// Taxa whose tip masks a wide-state workgroup stages in LDS.
constexpr int kMfmaTipLdsTaxa = 128;

// TL (n_taxa <= kMfmaTipLdsTaxa): the workgroup first copies the tip masks of
// its 16 G sites into LDS (tl[16 G t + site]).
// MP: the program-major launch of prune4_kernel (n_main site blocks, n_tail
// programs, blk0 of them a workgroup, grouped by XCD), one program after
// another in the workgroup.
template <typename T, int C, int V, int G, bool TL, bool MP>
__device__ __forceinline__ void prune_mfma_walk(
    const Op* __restrict__ ops, const int* __restrict__ prog_begin, const uint32_t* __restrict__ tips, T* clv,
    int* scl, const double* __restrict__ Pm, const double* __restrict__ PTm, const double* __restrict__ roots,
    const double* __restrict__ site_w, double* __restrict__ site_lnl, double* __restrict__ partial, int m, int m_pad,
    int S, uint32_t full, int n_taxa, int n_main, int prog_main, int blk0, int prog_piece, int tail0, int n_tail,
    unsigned* fold_counter, double* fold_out) {
    __shared__ int xsc[C][16 * G];
    __shared__ double xroot[C][16 * G];
    __shared__ uint32_t tl[TL ? kMfmaTipLdsTaxa * 16 * G : 1];
    __shared__ double xb[C][G * V * 64];    // each wave's carry B (LDS keeps the registers for the products)
    // work mapping as in prune4_kernel (plain or tail-split launch)
    const int y = blockIdx.y;
    int bx = blockIdx.x, prog = y;
    bool has_root = true;
    if (prog_main >= 0) {
        if ((int)blockIdx.x < n_main) {
            bx = blk0 + (int)blockIdx.x;
            prog = prog_main;
        } else {
            const int r = (int)blockIdx.x - n_main;
            bx = tail0 + r % n_tail;
            prog = prog_piece + r / n_tail;
            has_root = false;
        }
    } else if (MP) {
        const int slot = (int)blockIdx.x >> 3, n_pg = (n_tail + blk0 - 1) / blk0;
        prog = slot % n_pg * blk0;
        bx = slot / n_pg * 8 + ((int)blockIdx.x & 7);
        if (bx >= n_main) return;        // the last group's padding (uniform per workgroup)
    }
    if (TL) {
        const uint32_t* blk = tips + (size_t)bx * 16 * G;
        for (int i = threadIdx.x; i < n_taxa * 16 * G; i += 64 * C)
            tl[i] = blk[(size_t)(i / (16 * G)) * m_pad + i % (16 * G)];
        __syncthreads();
    }
    const int lane = threadIdx.x & 63;
    const int g = lane >> 4, j = lane & 15;
    const int cat = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int sg0 = bx * G, n_sg = m_pad >> 4;
    int s[G];
#pragma unroll
    for (int q = 0; q < G; ++q) s[q] = (sg0 + q) * 16 + j;
    const int p_end = MP ? min(prog + blk0, n_tail) : prog + 1;
    int p = prog;
    do {    // MP: the workgroup's programs in turn; otherwise once
    const int k0 = prog_begin[p], k1 = prog_begin[p + 1];
    const size_t slot_stride = (size_t)C * m_pad * 4 * V;
    // carried results stay in FP64 and are rescaled lazily, as in the DNA
    // walker; only stored partials are rounded to T
    double cA[G][V];
    int sA[G], sB[G];
    double* const cB = &xb[cat][lane];      // carry B: (group q, slot t) at cB[64 (q V + t)]
#pragma unroll
    for (int q = 0; q < G; ++q) {
#pragma unroll
        for (int t = 0; t < V; ++t) { cA[q][t] = 0.0; cB[(q * V + t) * 64] = 0.0; }
        sA[q] = 0;
        sB[q] = 0;
    }
    double lnl[G];
#pragma unroll
    for (int q = 0; q < G; ++q) lnl[q] = 0.0;
    Op nxt = ops[k0];
    for (int k = k0; k < k1; ++k) {
        const Op op = nxt;
        nxt = ops[k + 1];     // in flight during this op (an op buffer ends with one pad op)
        int sc[G];
#pragma unroll
        for (int q = 0; q < G; ++q) sc[q] = 0;
        double val[G][V];
#define PHYLO_MFMA_SHAPE(LK, RK)                                                                                     \
    mfma_shape<LK, RK, T, C, V, G, TL>(op, cat, s, sg0, n_sg, lane, g, m_pad, S, full, tips, tl, clv, slot_stride, \
                                       scl, Pm, PTm, cA, sA, cB, sB, val, sc)
        if (op.flags & OP_SH_AT) PHYLO_MFMA_SHAPE(SRC_CARRY_A, SRC_TIP);
        else if (op.flags & OP_SH_TT) PHYLO_MFMA_SHAPE(SRC_TIP, SRC_TIP);
        else if (op.flags & OP_SH_AB) PHYLO_MFMA_SHAPE(SRC_CARRY_A, SRC_CARRY_B);
        else if (op.flags & OP_SH_AC) PHYLO_MFMA_SHAPE(SRC_CARRY_A, SRC_CLV);
        else {
#undef PHYLO_MFMA_SHAPE
            // the other shapes: one child at a time, kinds read at run time
            mfma_d4 acc[G][2];
            mfma_child<T, C, V, G, TL>(op.l_kind, op.l_idx, op.l_node, cat, s, sg0, n_sg, lane, g, m_pad, S, full,
                                       tips, tl, clv, slot_stride, scl, Pm, PTm, cA, sA, cB, sB, acc, sc);
#pragma unroll
            for (int q = 0; q < G; ++q)
#pragma unroll
                for (int t = 0; t < V; ++t) val[q][t] = acc[q][t / 4][t % 4];
            if (op.r_kind != SRC_NONE) {
                mfma_child<T, C, V, G, TL>(op.r_kind, op.r_idx, op.r_node, cat, s, sg0, n_sg, lane, g, m_pad, S,
                                           full, tips, tl, clv, slot_stride, scl, Pm, PTm, cA, sA, cB, sB, acc, sc);
#pragma unroll
                for (int q = 0; q < G; ++q)
#pragma unroll
                    for (int t = 0; t < V; ++t) val[q][t] *= acc[q][t / 4][t % 4];
            }
        }
        // Lazy rescale: only when a site's largest state falls below 2^-256
        // (a site's states are spread over the lane groups, so its maximum
        // takes two cross-lane steps; every lane holding part of such a site
        // has its own largest state below the threshold, so those steps run
        // only when some lane of the wave has).
        bool low = false;
#pragma unroll
        for (int q = 0; q < G; ++q) {
            int h = __double2hiint(val[q][0]);     // the lane's largest state of the site, as a high word
#pragma unroll
            for (int t = 1; t < V; ++t) h = max(h, __double2hiint(val[q][t]));
            low |= h < kLazyRescaleHi;
        }
        if (__builtin_amdgcn_ballot_w64(low) != 0) {
#pragma unroll
            for (int q = 0; q < G; ++q) {
                const int e = site_exponent<V>(val[q]);
                const double mx = site_max<V>(val[q]);
                if (mx < 0x1p-256) {
                    sc[q] += e;
#pragma unroll
                    for (int t = 0; t < V; ++t) val[q][t] = ldexp(val[q][t], -e);
                }
            }
        }
#pragma unroll
        for (int q = 0; q < G; ++q) {
            if (op.out_slot >= 0) {
                // stored: brought to [0.5, 1) with its own scaler, then rounded
                // to T with per-state random bits
                const int e = site_exponent<V>(val[q]);
                const uint32_t key = round_key(s[q], op, cat) + 0x9e3779b9u * (8u * (uint32_t)g);
                T* dst = clv + (size_t)op.out_slot * slot_stride + ((size_t)cat * n_sg + sg0 + q) * V * 64 + lane;
#pragma unroll
                for (int t = 0; t < V; ++t)
                    dst[t * 64] = (T)round_cut(ldexp(val[q][t], -e), key + 0x9e3779b9u * t, T());
                if (g == 0) scl[((size_t)op.out_slot * C + cat) * m_pad + s[q]] = sc[q] + e;
            }
            if (op.flags & OP_TO_B) {
#pragma unroll
                for (int t = 0; t < V; ++t) cB[(q * V + t) * 64] = val[q][t];
                sB[q] = sc[q];
            } else {
#pragma unroll
                for (int t = 0; t < V; ++t) cA[q][t] = val[q][t];
                sA[q] = sc[q];
            }
        }
        if (op.flags & OP_ROOT) {
            double lc[G];
#pragma unroll
            for (int q = 0; q < G; ++q) {
                // the root partial brought to [0.5, 1) first (as in walk4)
                const int e = site_exponent<V>(val[q]);
                sc[q] += e;
#pragma unroll
                for (int t = 0; t < V; ++t) val[q][t] = ldexp(val[q][t], -e);
                lc[q] = 0.0;
#pragma unroll
                for (int t = 0; t < V; ++t) lc[q] = fma(roots[4 * t + g], val[q][t], lc[q]);
                lc[q] += __shfl_xor(lc[q], 16, 64);
                lc[q] += __shfl_xor(lc[q], 32, 64);
                lc[q] *= roots[kRootCatOff + cat];
            }
            if (C > 1) {
                // the categories' sums at their largest exponent
                if (g == 0) {
#pragma unroll
                    for (int q = 0; q < G; ++q) {
                        xroot[cat][16 * q + j] = lc[q];
                        xsc[cat][16 * q + j] = sc[q];
                    }
                }
                __syncthreads();
#pragma unroll
                for (int q = 0; q < G; ++q) {
                    int M = xsc[0][16 * q + j];
#pragma unroll
                    for (int c = 1; c < C; ++c) M = max(M, xsc[c][16 * q + j]);
                    lc[q] = 0.0;
#pragma unroll
                    for (int c = 0; c < C; ++c) lc[q] += ldexp(xroot[c][16 * q + j], xsc[c][16 * q + j] - M);
                    sc[q] = M;
                }
                __syncthreads();
            }
#pragma unroll
            for (int q = 0; q < G; ++q) lnl[q] = log(lc[q]) + sc[q] * kLn2;
        }
    }
    if (!has_root) return;    // a piece: its stores are its result (uniform per workgroup)
    double sum = 0.0;
#pragma unroll
    for (int q = 0; q < G; ++q) {
        const bool live = s[q] < m && cat == 0 && g == 0;
        if (live && site_lnl) site_lnl[(size_t)(MP ? p : y) * m + s[q]] = lnl[q];
        if (live) sum += (site_w ? site_w[s[q]] : 1.0) * lnl[q];
    }
    if constexpr (MP) {
        // only the category-0 wave holds site values: its sum is the block's
        if (cat == 0) {
            const double v = wave_sum(sum);
            if (lane == 0) partial[(size_t)p * n_main + bx] = v;
        }
    } else {
        block_sum_store(sum, partial + (prog_main >= 0 ? (size_t)bx : (size_t)y * gridDim.x + blockIdx.x));
        if (64 * C == kSumThreads && fold_counter) fold_total(partial, blk0 + n_main, fold_counter, fold_out);
    }
    } while (MP && ++p < p_end);
}

template <typename T, int C, int V, int G, bool TL>
__global__ __launch_bounds__(64 * C, (V <= 5 ? 3 : 1)) void prune_mfma_kernel(
    const Op* __restrict__ ops, const int* __restrict__ prog_begin, const uint32_t* __restrict__ tips, T* clv,
    int* scl, const double* __restrict__ Pm, const double* __restrict__ PTm, const double* __restrict__ roots,
    const double* __restrict__ site_w, double* __restrict__ site_lnl, double* __restrict__ partial, int m, int m_pad,
    int S, uint32_t full, int n_taxa, int n_main, int prog_main, int blk0, int prog_piece, int tail0, int n_tail,
    unsigned* fold_counter, double* fold_out) {
    prune_mfma_walk<T, C, V, G, TL, false>(ops, prog_begin, tips, clv, scl, Pm, PTm, roots, site_w, site_lnl, partial, m, m_pad, S, full,
                                          n_taxa, n_main, prog_main, blk0, prog_piece, tail0, n_tail, fold_counter,
                                          fold_out);
}

// An NNI neighbourhood's programs, program-major and grouped by XCD
// (prune_mfma_walk's MP launch).
template <typename T, int C, int V, int G>
__global__ __launch_bounds__(64 * C, (V <= 5 ? 3 : 1)) void prune_mfma_nni_kernel(
    const Op* __restrict__ ops, const int* __restrict__ prog_begin, const uint32_t* __restrict__ tips, T* clv,
    int* scl, const double* __restrict__ Pm, const double* __restrict__ PTm, const double* __restrict__ roots,
    const double* __restrict__ site_w, double* __restrict__ site_lnl, double* __restrict__ partial, int m, int m_pad,
    int S, uint32_t full, int n_taxa, int n_main, int prog_main, int blk0, int prog_piece, int tail0, int n_tail,
    unsigned* fold_counter, double* fold_out) {
    prune_mfma_walk<T, C, V, G, false, true>(ops, prog_begin, tips, clv, scl, Pm, PTm, roots, site_w, site_lnl, partial, m, m_pad, S, full,
                                          n_taxa, n_main, prog_main, blk0, prog_piece, tail0, n_tail, fold_counter,
                                          fold_out);
}

// Compensated fixed-order sum of the per-workgroup partials of each program
// (one workgroup per program, block_total).
__global__ __launch_bounds__(kSumThreads) void final_sum_kernel(const double* __restrict__ partial, int n_part,
                                                               double* __restrict__ out) {
    const double t = block_total(partial + (size_t)blockIdx.x * n_part, n_part);
    if (threadIdx.x == 0) out[blockIdx.x] = t;
}

// P[node][cat][i][j] = sum_k V[i][k] exp(lam_k r_c t_node) Vinv[k][j], padded to Sp x Sp with zeros.
__global__ void pbuild_eigen_kernel(const double* __restrict__ V, const double* __restrict__ Vinv,
                                    const double* __restrict__ lam, const double* __restrict__ eff,
                                    const double* __restrict__ cat_rate, int S, int Sp, int C, int n_nodes,
                                    double* __restrict__ P) {
    const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long total = (long)n_nodes * C * Sp * Sp;
    if (gid >= total) return;
    const int j = gid % Sp;
    const int i = (gid / Sp) % Sp;
    const int c = (gid / ((long)Sp * Sp)) % C;
    const int node = gid / ((long)Sp * Sp * C);
    double v = 0.0;
    if (i < S && j < S) {
        const double t = eff[node] * cat_rate[c];
        for (int k = 0; k < S; ++k) v = fma(V[i * S + k] * exp(lam[k] * t), Vinv[k * S + j], v);
    }
    P[gid] = v;
}

__device__ inline uint64_t sim_mix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
// Uniform in [0, 1) with 53 random bits from the counter (seed, site, stream).
__device__ inline double sim_uniform(uint64_t seed, uint64_t site, uint64_t stream) {
    const uint64_t h = sim_mix(seed ^ sim_mix(site * 0x9e3779b97f4a7c15ull ^ sim_mix(stream + 0x632be59bd9b4e019ull)));
    return (double)(h >> 11) * 0x1.0p-53;
}

// Forward simulation, one thread per site (simpleML::simulate_chars,
// simpleML.cpp:233-292): a rate category from the category weights, the root
// state from root_cum by the reference's subtract-until-non-positive walk
// (state 0 when it never gets there), then down the preorder each node's
// state from row P[parent state] the same way (the parent's state when the
// walk does not end).  States go to out[node][site].
__global__ __launch_bounds__(256) void simulate_kernel(const int* __restrict__ pre, const int* __restrict__ parent,
                                                       int n_nodes, int m, int S, int Sp, int C,
                                                       const double* __restrict__ P, const double* __restrict__ root_f,
                                                       const double* __restrict__ cat_w, uint64_t seed,
                                                       uint8_t* __restrict__ out) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= m) return;
    int cat = 0;
    if (C > 1) {
        double u = sim_uniform(seed, s, 0);
        cat = C - 1;
        for (int k = 0; k < C; ++k) {
            u -= cat_w[k];
            if (u < 0.0) { cat = k; break; }
        }
    }
    double u = sim_uniform(seed, s, 1);
    int st = 0;
    for (int j = 0; j < S; ++j) {
        u -= root_f[j];
        if (u <= 0.0) { st = j; break; }
    }
    out[(size_t)pre[0] * m + s] = (uint8_t)st;
    for (int k = 1; k < n_nodes; ++k) {
        const int v = pre[k];
        const int anc = out[(size_t)parent[v] * m + s];
        const double* row = P + (((size_t)v * C + cat) * Sp + anc) * Sp;
        double r = sim_uniform(seed, s, 2 + (uint64_t)v);
        int x = anc;
        for (int j = 0; j < S; ++j) {
            r -= row[j];
            if (r <= 0.0) { x = j; break; }
        }
        out[(size_t)v * m + s] = (uint8_t)x;
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

struct phylo_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string err;
    int prec = PHYLO_PREC_FP32;
    double last_ms = 0.0;

    // problem
    int n_taxa = 0, m = 0, S = 0, m_pad = 0;
    std::vector<uint32_t> masks;
    bool have_tips = false, dirty_tips = true;
    std::vector<double> site_w;
    bool have_site_w = false, dirty_site_w = false;
    int n_nodes = 0, root = -1;
    std::vector<int> left, right;
    bool have_topo = false, dirty_topo = true;
    std::vector<double> eff;
    bool have_eff = false;
    int S_model = 0, builder = PHYLO_P_EIGEN;
    std::vector<double> Q, root_w, pi, V, Vinv, lam;
    bool have_model = false;
    int C = 1;
    std::vector<double> cat_rate{1.0}, cat_w{1.0};
    bool dirty_P = true;
    bool dirty_roots = true;
    double base_total = 0.0;
    long nni_ops_last = 0;

    // compiled program
    std::vector<Op> ops;
    std::vector<int> prog_begin;
    int n_slots = 0;
    // tail split of whole-tree DNA walks (build_split_programs): split_K chunk
    // programs from program id split_base[full ? 0 : 1], then one top program
    int split_K = 0, split_base[2] = {0, 0};
    bool split_all = false;                               // every block walks the pieces (too few blocks)
    bool tail_split = true;                               // phylo_set_tail_split
    int n_cu = 0;                                         // compute units of the device
    bool run_full = false;                                // the whole-tree launch stores every partial
    bool folded = false;                                  // the last launch wrote the total itself (fold_total)
    unsigned* d_fold = nullptr;                           // fold_total's counter (zero between launches)

    // device buffers
    void* d_tips = nullptr; size_t tips_bytes = 0;     // uint8 codes (S <= 4) or uint32 masks
    double* d_Pm = nullptr; size_t Pm_bytes = 0;       // P in MFMA operand order (wide states)
    double* d_PTm = nullptr; size_t PTm_bytes = 0;     // P columns for tips (wide states)
    void* d_clv = nullptr; size_t clv_bytes = 0;
    int* d_scl = nullptr; size_t scl_bytes = 0;
    double* d_P = nullptr; size_t P_bytes = 0;
    Op* d_ops = nullptr; size_t ops_bytes = 0;
    int* d_prog = nullptr; size_t prog_bytes = 0;
    Op* d_nni_ops = nullptr; size_t nni_ops_bytes = 0;   // NNI partial-recompute programs
    int* d_nni_prog = nullptr; size_t nni_prog_bytes = 0;
    const Op* run_ops = nullptr;                          // program set of the launch in flight
    const int* run_prog = nullptr;
    bool run_whole_tree = false;                          // the launch in flight walks every internal node
    bool clv_valid = false;                               // every stored partial matches the current inputs
    bool keep_partials = false;                           // loglik stores every node's partial
    double* d_roots = nullptr;
    double* d_sitew = nullptr; size_t sitew_bytes = 0;
    double* d_sitelnl = nullptr; size_t sitelnl_bytes = 0;
    double* d_partial = nullptr; size_t partial_bytes = 0;
    double* d_eig = nullptr; size_t eig_bytes = 0;      // V, Vinv, lambda, category rates, edge lengths
    double* h_total = nullptr; size_t h_total_bytes = 0;  // pinned, mapped: final_sum_kernel writes here
    double* h_total_dev = nullptr;                        // its device address
    // Pinned staging for the per-evaluation uploads (eigen system + edge
    // lengths, root / category weights): copied without a stream
    // synchronisation; an event guards reuse of the host buffer.
    struct Stage {
        double* h = nullptr;
        size_t cap = 0;
        hipEvent_t ev = nullptr;
        bool pending = false;
    } stage_P, stage_roots;
};

namespace {

int fail(phylo_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

#define HIPCHK(ctx, expr)                                                                   \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return fail((ctx), PHYLO_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// (Re)allocates a device buffer when it has to grow.
template <typename T>
hipError_t ensure(T*& ptr, size_t& have, size_t need) {
    if (need <= have && ptr) return hipSuccess;
    if (ptr) { hipError_t e = hipFree(ptr); if (e != hipSuccess) return e; ptr = nullptr; have = 0; }
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&ptr), need ? need : 16);
    if (e == hipSuccess) have = need;
    return e;
}

// Puts each op's children in kind order (carry A, tip, carry B, stored
// partial, none), swapping them where needed and marking the op OP_SWAP, so
// the walkers meet few child shapes (a Yule tree's ops take four) and the
// DNA walker's products start with the child already in registers while a
// tip's LDS reads are in flight.  No value changes: the children's products
// are multiplied elementwise, which commutes, and round_key undoes the swap.
// Idempotent.
void order_children(std::vector<Op>& ops) {
    static const int rank[5] = {1, 3, 0, 2, 4};    // by SrcKind: TIP, CLV, CARRY_A, CARRY_B, NONE
    for (Op& o : ops) {
        if (rank[o.r_kind] < rank[o.l_kind]) {
            std::swap(o.l_node, o.r_node);
            std::swap(o.l_kind, o.r_kind);
            std::swap(o.l_idx, o.r_idx);
            o.flags ^= OP_SWAP;
        }
        o.flags &= ~OP_SH_ALL;
        if (o.l_kind == SRC_CARRY_A && o.r_kind == SRC_TIP) o.flags |= OP_SH_AT;
        if (o.l_kind == SRC_TIP && o.r_kind == SRC_TIP) o.flags |= OP_SH_TT;
        if (o.l_kind == SRC_CARRY_A && o.r_kind == SRC_CARRY_B) o.flags |= OP_SH_AB;
        if (o.l_kind == SRC_CARRY_A && o.r_kind == SRC_CLV) o.flags |= OP_SH_AC;
    }
}

// Copies a program's ops to the device, children in kind order, followed by
// one zero op: the walkers read the op after the one they run (up to
// ops[end] of the last program), the DNA walker to load its P matrices
// early, the matrix walker to have the descriptor at hand.
hipError_t upload_ops(Op*& d, size_t& have, std::vector<Op>& ops, hipStream_t st) {
    static const Op pad{};
    order_children(ops);
    hipError_t e = ensure(d, have, (ops.size() + 1) * sizeof(Op));
    if (e == hipSuccess) e = hipMemcpyAsync(d, ops.data(), ops.size() * sizeof(Op), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d + ops.size(), &pad, sizeof(Op), hipMemcpyHostToDevice, st);
    return e;
}

// Host staging area of n doubles, free to overwrite once the copy that last
// read it has completed.
hipError_t stage_acquire(phylo_ctx::Stage& st, size_t n, double** out) {
    hipError_t e;
    if (st.pending) {
        if ((e = hipEventSynchronize(st.ev)) != hipSuccess) return e;
        st.pending = false;
    }
    if (!st.ev && (e = hipEventCreateWithFlags(&st.ev, hipEventDisableTiming)) != hipSuccess) return e;
    if (st.cap < n) {
        if (st.h && (e = hipHostFree(st.h)) != hipSuccess) return e;
        st.h = nullptr;
        st.cap = 0;
        if ((e = hipHostMalloc(reinterpret_cast<void**>(&st.h), n * sizeof(double))) != hipSuccess) return e;
        st.cap = n;
    }
    *out = st.h;
    return hipSuccess;
}

hipError_t stage_release(phylo_ctx::Stage& st, hipStream_t stream) {
    hipError_t e = hipEventRecord(st.ev, stream);
    st.pending = e == hipSuccess;
    return e;
}

void stage_free(phylo_ctx::Stage& st) {
    if (st.pending && st.ev) (void)hipEventSynchronize(st.ev);
    if (st.h) (void)hipHostFree(st.h);
    if (st.ev) (void)hipEventDestroy(st.ev);
    st = phylo_ctx::Stage{};
}

// States per stored partial: 4 for the DNA walker, else the next supported
// multiple of 4 (padded rows/columns of P are zero).
int padded_states(int S) {
    if (S <= 4) return 4;
    if (S <= 8) return 8;
    if (S <= 16) return 16;
    if (S <= 20) return 20;
    if (S <= 24) return 24;
    return 32;
}

// States per stored partial site: the MFMA walker (S > 4) keeps 4 * V operand
// slots per site, the DNA walker 4.
int clv_states(const phylo_ctx* c) { return c->S > 4 ? 4 * mfma_slots(c->S) : 4; }
bool uses_mfma(const phylo_ctx* c) { return c->S > 4; }

// Postorder of the internal nodes below `root`, visiting first[v] before
// second[v] (either may be -1).  Returns the nodes visited.
template <class TV>
std::vector<int> internal_postorder(const TV* c, const std::vector<int>& first, const std::vector<int>& second) {
    std::vector<int> order;
    order.reserve(c->n_nodes);
    std::vector<int> stack{c->root};
    std::vector<char> seen(c->n_nodes, 0);
    while (!stack.empty()) {
        const int v = stack.back();
        if (!seen[v]) {
            seen[v] = 1;
            if (second[v] >= 0) stack.push_back(second[v]);
            if (first[v] >= 0) stack.push_back(first[v]);
        } else {
            stack.pop_back();
            if (v >= c->n_taxa) order.push_back(v);
        }
    }
    return order;
}

// Appends the full-evaluation program of a tree: every internal node once,
// children before parents, each op's result kept in a register for the next
// op (carry A).  A node with two internal children needs one child's result
// while the other child's subtree is walked; that result waits in a second
// register (carry B) when the other subtree has no such node itself, and in
// memory otherwise.  The subtree without branching goes second where one
// exists, so that as many waits as possible use the register: with two
// registers a 128-taxon Yule tree keeps 11 of 127 partials in memory instead
// of 40.  The order of the walk does not change any value (each op computes
// (P_l L_l) x (P_r L_r) of the same children); the reference's recursion is
// left subtree, right subtree, node (simpleML.cpp:127-184).
// Returns false if internal nodes are unreachable.
template <class TV>
bool build_full_ops(const TV* c, std::vector<Op>& ops) {
    const int n = c->n_nodes, nt = c->n_taxa;
    std::vector<int> first(c->left.begin(), c->left.begin() + n), second(c->right.begin(), c->right.begin() + n);
    const std::vector<int> post = internal_postorder(c, first, second);
    if ((int)post.size() != n - nt) return false;
    const auto internal = [&](int x) { return x >= nt; };
    // branchy[v]: v's subtree holds a node with two internal children
    std::vector<char> branchy(n, 0), to_b(n, 0);
    for (int v : post) {
        const int l = c->left[v], r = c->right[v];
        const bool two = internal(l) && r >= 0 && internal(r);
        branchy[v] = two || (l >= 0 && branchy[l]) || (r >= 0 && branchy[r]);
        if (two) {
            if (branchy[r] && !branchy[l]) std::swap(first[v], second[v]);
            to_b[first[v]] = !branchy[second[v]];
        }
    }
    const std::vector<int> order = internal_postorder(c, first, second);
    int prev = -1, in_b = -1;     // node whose result is in carry A / carry B
    auto src = [&](int child, int& kind, int& idx) {
        if (child < 0) { kind = SRC_NONE; idx = 0; }
        else if (child < nt) { kind = SRC_TIP; idx = child; }
        else if (child == prev) { kind = SRC_CARRY_A; idx = 0; }
        else if (child == in_b) { kind = SRC_CARRY_B; idx = 0; }
        else { kind = SRC_CLV; idx = child - nt; }
    };
    for (int v : order) {
        Op op{};
        op.l_node = c->left[v];
        op.r_node = c->right[v] < 0 ? 0 : c->right[v];
        src(c->left[v], op.l_kind, op.l_idx);
        src(c->right[v], op.r_kind, op.r_idx);
        if (op.l_kind == SRC_CARRY_B || op.r_kind == SRC_CARRY_B) in_b = -1;
        op.out_slot = v - nt;
        op.flags = (v == c->root) ? OP_ROOT : 0;
        if (to_b[v] && v != c->root) {
            op.flags |= OP_TO_B;
            in_b = v;
            prev = -1;
        } else {
            prev = v;
        }
        ops.push_back(op);
    }
    return true;
}

// The DNA walker reads a program's root value from carry A after its last
// op, so a root op must end its program.  Every program builder puts the root
// last (children before parents); this checks it before programs go to the
// device.
bool root_ops_last(const std::vector<Op>& ops, const std::vector<int>& begin) {
    for (size_t p = 0; p + 1 < begin.size(); ++p)
        for (int k = begin[p]; k < begin[p + 1]; ++k)
            if ((ops[k].flags & OP_ROOT) && k != begin[p + 1] - 1) return false;
    return true;
}

// Clears the stores of ops[begin, end) whose result no op of that range reads
// back from memory: results consumed only through the two carries.  With the
// second carry only results that wait while a branching subtree is walked keep
// a store (11 of the 127 internal nodes of a 128-taxon Yule tree), so the
// partials written per evaluation drop by more than 90%.
void drop_unread_stores(std::vector<Op>& ops, size_t begin, size_t end) {
    std::vector<char> read;
    for (size_t k = begin; k < end; ++k) {
        const Op& o = ops[k];
        for (int side = 0; side < 2; ++side) {
            const int kind = side ? o.r_kind : o.l_kind, idx = side ? o.r_idx : o.l_idx;
            if (kind != SRC_CLV) continue;
            if ((size_t)idx >= read.size()) read.resize(idx + 1, 0);
            read[idx] = 1;
        }
    }
    for (size_t k = begin; k < end; ++k) {
        Op& o = ops[k];
        if (o.out_slot >= 0 && ((size_t)o.out_slot >= read.size() || !read[o.out_slot])) o.out_slot = -1;
    }
}

// Two programs over the same postorder: 0 stores every internal node's
// partial (needed by node_partials and the NNI partial recompute), 1 stores
// only what the walk itself reads back.  Both compute identical values.
// The tail split of a whole-tree DNA walk.  When the 64-site blocks do not
// divide evenly among the CUs, the blocks past the last full round (the
// tail) put one more workgroup on some CUs, and those CUs set the time of the
// whole launch (128-taxon GTR+G4 tree: 1,536 blocks on 256 CUs 0.128 ms,
// 1,564 blocks 0.162 ms).  The tail blocks instead walk the tree in pieces:
// the tree is cut into subtrees of at most L internal nodes, packed in
// postorder into chunks of at most L ops whose walks run as extra workgroups
// beside the main blocks and store their results; a second launch then walks
// the few nodes above the cut (the top program) for the tail blocks only.
// Chunk and top programs use carry A alone: a result is stored unless the
// next op reads it (every result in the storing variant).  Appends, for the
// storing and the lean evaluation, split_K chunk programs and one top program
// to c->ops / c->prog_begin; split_K stays 0 when no cut keeps the top
// program short (a caterpillar tree, say) or the tree is small.
bool build_split_programs(phylo_ctx* c, int max_chunks, bool split_all) {
    c->split_K = 0;
    const int nt = c->n_taxa, nn = c->n_nodes, root = c->root;
    const std::vector<int> post = internal_postorder(c, c->left, c->right);
    const int ni = (int)post.size();
    if (ni < 16) return false;
    std::vector<int> size(nn, 0), par(nn, -1);
    for (int i = 0; i < ni; ++i) {
        const int v = post[i], l = c->left[v], r = c->right[v];
        size[v] = 1 + (l >= nt ? size[l] : 0) + (r >= nt ? size[r] : 0);
        if (l >= 0) par[l] = v;
        if (r >= 0) par[r] = v;
    }
    // pieces: the largest subtrees of at most L internal nodes, in postorder;
    // a subtree's internal nodes are the size[v] entries of post ending at v
    const auto cut = [&](int L, std::vector<std::vector<int>>& chunks, std::vector<int>& top) {
        chunks.clear();
        top.clear();
        std::vector<int> cur;
        int cur_n = 0;
        for (int i = 0; i < ni; ++i) {
            const int v = post[i];
            if (size[v] > L) { top.push_back(v); continue; }
            if (v != root && size[par[v]] <= L) continue;     // inside a larger piece
            if (cur_n > 0 && cur_n + size[v] > L) { chunks.push_back(cur); cur.clear(); cur_n = 0; }
            for (int k = i - size[v] + 1; k <= i; ++k) cur.push_back(post[k]);
            cur_n += size[v];
        }
        if (cur_n > 0) chunks.push_back(cur);
    };
    int best_L = -1;
    double best_cost = 1e30;
    std::vector<std::vector<int>> chunks;
    std::vector<int> top;
    // piece sizes to try: fractions of the tree, and (so that few large
    // pieces are possible) the sizes of the subtrees near the root
    std::vector<int> Ls;
    for (int d : {16, 12, 10, 8, 6, 5, 4, 3, 2}) Ls.push_back(std::max(1, ni / d));
    for (int v : {c->left[root], c->right[root]})
        if (v >= nt) {
            Ls.push_back(size[v]);
            for (int w : {c->left[v], c->right[v]})
                if (w >= nt) Ls.push_back(size[w]);
        }
    for (int L : Ls) {
        if (size[root] <= L) continue;
        cut(L, chunks, top);
        size_t big = 0;
        for (const auto& ch : chunks) big = std::max(big, ch.size());
        if (chunks.size() < 2 || top.size() > 16 || (int)chunks.size() > max_chunks) continue;
        // Tail split: the top program runs alone, about 0.7 us per op; a
        // chunk shares a CU with main blocks and costs nothing until it is
        // longer than about 0.4 of the whole walk, then about 0.5 us per op
        // (C2 tree, 127 ops, kernel span: top 1 / longest chunk 70 ops
        // 0.1134 ms, 2 / 56 0.1066, 3 / 49 0.1062, 4 / 44 0.1063, 7 / 34
        // 0.1074, 8 / 28 0.1083, 9 / 20 0.1089).  Every block split: the
        // longest chunk and then the top program form the chain.
        const double over = std::max(0.0, (double)big - 0.4 * ni);
        const double cost = split_all ? (double)big + (double)top.size() : (double)top.size() + 0.7 * over;
        if (cost < best_cost) { best_cost = cost; best_L = L; }
    }
    if (best_L < 0) return false;
    cut(best_L, chunks, top);
    const auto emit = [&](const std::vector<int>& nodes, bool store_all) {
        for (size_t i = 0; i < nodes.size(); ++i) {
            const int v = nodes[i], l = c->left[v], r = c->right[v];
            const auto src = [&](int ch, int& kind, int& idx) {
                if (ch < 0) { kind = SRC_NONE; idx = 0; }
                else if (ch < nt) { kind = SRC_TIP; idx = ch; }
                else if (i > 0 && nodes[i - 1] == ch) { kind = SRC_CARRY_A; idx = 0; }
                else { kind = SRC_CLV; idx = ch - nt; }
            };
            Op op{};
            op.l_node = l;
            op.r_node = r < 0 ? 0 : r;
            src(l, op.l_kind, op.l_idx);
            src(r, op.r_kind, op.r_idx);
            const bool next_reads = i + 1 < nodes.size() && par[v] == nodes[i + 1];
            op.out_slot = (store_all || !next_reads) ? v - nt : -1;
            op.flags = v == root ? OP_ROOT : 0;
            c->ops.push_back(op);
        }
        c->prog_begin.push_back((int)c->ops.size());
    };
    c->split_K = (int)chunks.size();
    for (int mode = 0; mode < 2; ++mode) {     // 0: store every partial, 1: lean
        c->split_base[mode] = (int)c->prog_begin.size() - 1;
        for (const auto& ch : chunks) emit(ch, mode == 0);
        emit(top, mode == 0);
    }
    return true;
}

// Workgroups of the whole-tree walker that one CU holds at once, for this
// context's alphabet, precision and rate categories (the occupancy API on the
// kernel instance the launch uses); 0 if unknown.
template <typename T, int C>
int resident_dna(const phylo_ctx* c) {
    int n = 0;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, prune4_kernel<T, C, true>, 64 * C,
                                                        tip_lds_bytes(c->n_taxa)) == hipSuccess ? n : 0;
}
template <typename T, int C, int V>
int resident_wide(const phylo_ctx*) {
    int n = 0;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, prune_mfma_kernel<T, C, V, kMfmaGroups, true>, 64 * C, 0) ==
                   hipSuccess ? n : 0;
}
template <typename T, int C>
int resident_wide_v(const phylo_ctx* c) {
    switch (mfma_slots(c->S)) {
    case 2: return resident_wide<T, C, 2>(c);
    case 3: return resident_wide<T, C, 3>(c);
    case 4: return resident_wide<T, C, 4>(c);
    case 5: return resident_wide<T, C, 5>(c);
    case 6: return resident_wide<T, C, 6>(c);
    case 7: return resident_wide<T, C, 7>(c);
    default: return resident_wide<T, C, 8>(c);
    }
}
template <typename T>
int resident_per_cu_t(const phylo_ctx* c) {
    const bool wide = c->S > 4;
    switch (c->C) {
    case 1: return wide ? resident_wide_v<T, 1>(c) : resident_dna<T, 1>(c);
    case 2: return wide ? resident_wide_v<T, 2>(c) : resident_dna<T, 2>(c);
    case 4: return wide ? resident_wide_v<T, 4>(c) : resident_dna<T, 4>(c);
    case 8: return wide ? resident_wide_v<T, 8>(c) : resident_dna<T, 8>(c);
    default: return 0;
    }
}
int resident_per_cu(const phylo_ctx* c) {
    return c->prec == PHYLO_PREC_FP64 ? resident_per_cu_t<double>(c) : resident_per_cu_t<float>(c);
}

int compile_full_program(phylo_ctx* c) {
    c->ops.clear();
    if (!build_full_ops(c, c->ops)) return fail(c, PHYLO_ERR_ARG, "topology: internal nodes unreachable from the root");
    const size_t n = c->ops.size();
    c->ops.insert(c->ops.end(), c->ops.begin(), c->ops.end());
    drop_unread_stores(c->ops, n, 2 * n);
    c->prog_begin = {0, (int)n, (int)(2 * n)};
    c->n_slots = c->n_nodes - c->n_taxa;
    // When every site block is split (too few blocks to fill the CUs), its
    // pieces run as K workgroups per block: K is capped so that all of them
    // are resident at once, or the pieces would run in rounds and the walk
    // would take no less than the whole-tree chain.  A tail beside full
    // rounds needs no cap.
    c->split_all = false;
    const int nblk = c->m_pad / (uses_mfma(c) ? 16 * kMfmaGroups : 64);
    if (c->n_cu > 0 && nblk < kSplitAllBelowPerCu * c->n_cu) {
        const int per_cu = resident_per_cu(c);
        const int max_chunks = per_cu > 0 ? per_cu * c->n_cu / nblk : 0;
        const size_t n_ops = c->ops.size();
        const size_t n_prog = c->prog_begin.size();
        c->split_all = build_split_programs(c, max_chunks, true);
        if (!c->split_all) {                  // no such cut: the tail split
            c->ops.resize(n_ops);
            c->prog_begin.resize(n_prog);
        }
    }
    if (!c->split_all) build_split_programs(c, INT_MAX, false);
    return PHYLO_OK;
}

// Host-built P for REFCOMPAT / PADE, laid out [node][cat][Sp][Sp].
void host_pmatrices(const phylo_ctx* c, int Sp, std::vector<double>& P) {
    const int S = c->S_model;
    P.assign((size_t)c->n_nodes * c->C * Sp * Sp, 0.0);
    std::vector<double> tmp((size_t)S * S);
    for (int v = 0; v < c->n_nodes; ++v) {
        if (v == c->root) continue;
        for (int k = 0; k < c->C; ++k) {
            const double t = c->eff[v] * c->cat_rate[k];
            if (c->builder == PHYLO_P_REFCOMPAT) phylo_math::p_refcompat(S, c->Q.data(), t, tmp.data());
            else phylo_math::expm_pade13(S, c->Q.data(), t, tmp.data());
            double* dst = P.data() + ((size_t)v * c->C + k) * Sp * Sp;
            for (int i = 0; i < S; ++i)
                for (int j = 0; j < S; ++j) dst[i * Sp + j] = tmp[(size_t)i * S + j];
        }
    }
}

int upload_pmatrices(phylo_ctx* c) {
    const int Sp = padded_states(c->S_model);
    const size_t nP = (size_t)c->n_nodes * c->C * Sp * Sp;
    HIPCHK(c, ensure(c->d_P, c->P_bytes, nP * sizeof(double)));
    if (c->builder == PHYLO_P_EIGEN) {
        const int S = c->S_model;
        const size_t n_eig = 2 * c->V.size() + c->lam.size() + c->cat_rate.size();
        double* h = nullptr;
        HIPCHK(c, stage_acquire(c->stage_P, n_eig + c->eff.size(), &h));
        double* w = std::copy(c->V.begin(), c->V.end(), h);
        w = std::copy(c->Vinv.begin(), c->Vinv.end(), w);
        w = std::copy(c->lam.begin(), c->lam.end(), w);
        w = std::copy(c->cat_rate.begin(), c->cat_rate.end(), w);
        std::copy(c->eff.begin(), c->eff.end(), w);
        // eigen system and edge lengths in one buffer: one copy per model or length change
        HIPCHK(c, ensure(c->d_eig, c->eig_bytes, (n_eig + c->eff.size()) * sizeof(double)));
        HIPCHK(c, hipMemcpyAsync(c->d_eig, h, (n_eig + c->eff.size()) * sizeof(double), hipMemcpyHostToDevice,
                                 c->stream));
        const double* d_eff = c->d_eig + n_eig;
        HIPCHK(c, stage_release(c->stage_P, c->stream));
        const double* dV = c->d_eig;
        const double* dVi = dV + (size_t)S * S;
        const double* dl = dVi + (size_t)S * S;
        const double* dr = dl + S;
        const int threads = 256;
        const long blocks = ((long)nP + threads - 1) / threads;
        hipLaunchKernelGGL(pbuild_eigen_kernel, dim3((unsigned)blocks), dim3(threads), 0, c->stream, dV, dVi, dl,
                           d_eff, dr, S, Sp, c->C, c->n_nodes, c->d_P);
        HIPCHK(c, hipGetLastError());
    } else {
        std::vector<double> P;
        host_pmatrices(c, Sp, P);
        HIPCHK(c, hipMemcpyAsync(c->d_P, P.data(), P.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    if (Sp > 4) {
        const int V = mfma_slots(c->S_model), TN = V > 4 ? 2 : 1;
        const long nPm = (long)c->n_nodes * c->C * TN * V * 64;
        HIPCHK(c, ensure(c->d_Pm, c->Pm_bytes, (size_t)nPm * sizeof(double)));
        hipLaunchKernelGGL(pmfma_kernel, dim3((unsigned)((nPm + 255) / 256)), dim3(256), 0, c->stream, c->d_P,
                           c->S_model, Sp, c->C, c->n_nodes, V, TN, c->d_Pm);
        HIPCHK(c, hipGetLastError());
        const long nPT = (long)c->n_nodes * c->C * (c->S_model + 1) * 4 * V;
        HIPCHK(c, ensure(c->d_PTm, c->PTm_bytes, (size_t)nPT * sizeof(double)));
        hipLaunchKernelGGL(ptm_kernel, dim3((unsigned)((nPT + 255) / 256)), dim3(256), 0, c->stream, c->d_P,
                           c->S_model, Sp, c->C, c->n_nodes, 4 * V, c->d_PTm);
        HIPCHK(c, hipGetLastError());
    }
    c->dirty_P = false;
    return PHYLO_OK;
}

int prepare(phylo_ctx* c) {
    if (!c->have_tips) return fail(c, PHYLO_ERR_STATE, "phylo_set_tips has not been called");
    if (!c->have_topo) return fail(c, PHYLO_ERR_STATE, "phylo_set_topology has not been called");
    if (!c->have_eff) return fail(c, PHYLO_ERR_STATE, "phylo_set_edge_lengths has not been called");
    if (!c->have_model) return fail(c, PHYLO_ERR_STATE, "phylo_set_model has not been called");
    if (c->S_model != c->S) return fail(c, PHYLO_ERR_ARG, "model and tip data disagree on the number of states");
    HIPCHK(c, hipSetDevice(c->device));
    if (c->dirty_tips) {
        // [taxon][m_pad]; padding sites are all-states (finite, ignored in the sum)
        if (c->S <= 4) {
            std::vector<uint8_t> codes((size_t)c->n_taxa * c->m_pad, 0x0F);
            for (int t = 0; t < c->n_taxa; ++t)
                for (int s = 0; s < c->m; ++s)
                    codes[(size_t)t * c->m_pad + s] = (uint8_t)(c->masks[(size_t)t * c->m + s] & 0x0Fu);
            HIPCHK(c, ensure(c->d_tips, c->tips_bytes, codes.size()));
            HIPCHK(c, hipMemcpyAsync(c->d_tips, codes.data(), codes.size(), hipMemcpyHostToDevice, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
        } else {
            const uint32_t full = c->S == 32 ? 0xFFFFFFFFu : ((1u << c->S) - 1u);
            std::vector<uint32_t> codes((size_t)c->n_taxa * c->m_pad, full);
            for (int t = 0; t < c->n_taxa; ++t)
                for (int s = 0; s < c->m; ++s) codes[(size_t)t * c->m_pad + s] = c->masks[(size_t)t * c->m + s];
            HIPCHK(c, ensure(c->d_tips, c->tips_bytes, codes.size() * sizeof(uint32_t)));
            HIPCHK(c, hipMemcpyAsync(c->d_tips, codes.data(), codes.size() * sizeof(uint32_t), hipMemcpyHostToDevice,
                                     c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
        }
        c->dirty_tips = false;
    }
    if (c->dirty_site_w) {
        if (c->have_site_w) {
            HIPCHK(c, ensure(c->d_sitew, c->sitew_bytes, (size_t)c->m * sizeof(double)));
            HIPCHK(c, hipMemcpyAsync(c->d_sitew, c->site_w.data(), (size_t)c->m * sizeof(double), hipMemcpyHostToDevice, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
        }
        c->dirty_site_w = false;
    }
    if (c->dirty_topo) {
        int rc = compile_full_program(c);
        if (rc) return rc;
        if (!root_ops_last(c->ops, c->prog_begin)) return fail(c, PHYLO_ERR_STATE, "internal: root op not last");
        HIPCHK(c, upload_ops(c->d_ops, c->ops_bytes, c->ops, c->stream));
        HIPCHK(c, ensure(c->d_prog, c->prog_bytes, c->prog_begin.size() * sizeof(int)));
        HIPCHK(c, hipMemcpyAsync(c->d_prog, c->prog_begin.data(), c->prog_begin.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        c->dirty_topo = false;
    }
    const size_t elt = c->prec == PHYLO_PREC_FP64 ? sizeof(double) : sizeof(float);
    HIPCHK(c, ensure(c->d_clv, c->clv_bytes,
                     (size_t)std::max(c->n_slots, 1) * c->C * c->m_pad * clv_states(c) * elt));
    // scalers: [slot][cat][site]
    HIPCHK(c, ensure(c->d_scl, c->scl_bytes, (size_t)std::max(c->n_slots, 1) * c->C * c->m_pad * sizeof(int)));
    if (c->dirty_P) {
        int rc = upload_pmatrices(c);
        if (rc) return rc;
    }
    if (c->dirty_roots) {
        const size_t n_roots = kMaxStates + kMaxCat;
        double* roots = nullptr;
        HIPCHK(c, stage_acquire(c->stage_roots, n_roots, &roots));
        std::fill(roots, roots + n_roots, 0.0);
        for (int i = 0; i < c->S; ++i) roots[i] = c->root_w[i];
        for (int k = 0; k < c->C; ++k) roots[kRootCatOff + k] = c->cat_w[k];
        if (!c->d_roots) HIPCHK(c, hipMalloc(&c->d_roots, n_roots * sizeof(double)));
        HIPCHK(c, hipMemcpyAsync(c->d_roots, roots, n_roots * sizeof(double), hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, stage_release(c->stage_roots, c->stream));
        c->dirty_roots = false;
    }
    return PHYLO_OK;
}


// Site blocks of a whole-tree launch walked as whole-tree workgroups; the
// rest (the tail) go through the split programs.  Full rounds over the CUs
// stay whole unless there are too few of them.
int split_main_blocks(const phylo_ctx* c, int nblk) {
    if (c->n_cu <= 0) return nblk;
    if (c->split_all) return 0;
    return nblk / c->n_cu * c->n_cu;
}

template <typename T, int C>
void launch_prune4(phylo_ctx* c, dim3 grid, double* site_lnl) {
    const auto* tips = reinterpret_cast<const uint8_t*>(c->d_tips);
    T* clv = reinterpret_cast<T*>(c->d_clv);
    const double* sw = c->have_site_w ? c->d_sitew : nullptr;
    // staging every tip pays off for whole-tree walks, not for the short
    // path-to-root programs of an NNI neighbourhood
    if (c->run_whole_tree && c->n_taxa <= kTipLdsTaxa) {
        const int nblk = (int)grid.x;
        const int main_b = split_main_blocks(c, nblk), tail = nblk - main_b;
        if (c->tail_split && c->split_K > 0 && grid.y == 1 && tail > 0) {
            // the main blocks and the tail's pieces, then the tail's top program
            const int base = c->split_base[c->run_full ? 0 : 1], K = c->split_K;
            hipLaunchKernelGGL((prune4_kernel<T, C, true>), dim3(main_b + K * tail), dim3(64 * C),
                               tip_lds_bytes(c->n_taxa), c->stream,
                               c->run_ops, c->d_prog, tips, clv, c->d_scl, c->d_P, c->d_roots, sw, site_lnl,
                               c->d_partial, c->m, c->m_pad, c->n_taxa, main_b, c->run_full ? 0 : 1, 0, base,
                               main_b, tail, nullptr, nullptr);
            // (the top program reads a few tips: no LDS staging)
            unsigned* fold = 64 * C == kSumThreads ? c->d_fold : nullptr;
            hipLaunchKernelGGL((prune4_kernel<T, C, false>), dim3(tail), dim3(64 * C), 0, c->stream, c->run_ops,
                               c->d_prog, tips, clv, c->d_scl, c->d_P, c->d_roots, sw, site_lnl, c->d_partial, c->m,
                               c->m_pad, c->n_taxa, tail, base + K, main_b, 0, 0, 1, fold, c->h_total_dev);
            c->folded = fold != nullptr;
        } else {
            hipLaunchKernelGGL((prune4_kernel<T, C, true>), grid, dim3(64 * C), tip_lds_bytes(c->n_taxa), c->stream,
                               c->run_ops, c->run_prog,
                               tips, clv, c->d_scl, c->d_P, c->d_roots, sw, site_lnl, c->d_partial, c->m, c->m_pad,
                               c->n_taxa, (int)grid.x, -1, 0, 0, 0, 1, nullptr, nullptr);
        }
    } else if (grid.y > 1) {
        // many programs (an NNI neighbourhood): program-major, grouped by
        // XCD, four programs a workgroup while that leaves 8 workgroups per CU
        const int nblk = (int)grid.x, n_prog = (int)grid.y, nb8 = (nblk + 7) / 8 * 8;
        const long fill = 8L * (c->n_cu > 0 ? c->n_cu : 256);
        int ppw = 4;
        while (ppw > 1 && (long)nb8 * ((n_prog + ppw - 1) / ppw) < fill) ppw /= 2;
        const int n_pg = (n_prog + ppw - 1) / ppw;
        hipLaunchKernelGGL((prune4_kernel<T, C, false>), dim3(nb8 * n_pg), dim3(64 * C), 0,
                           c->stream, c->run_ops, c->run_prog, tips, clv, c->d_scl, c->d_P, c->d_roots, sw, site_lnl,
                           c->d_partial, c->m, c->m_pad, c->n_taxa, nblk, -2, ppw, 0, 0, n_prog, nullptr, nullptr);
    } else {
        hipLaunchKernelGGL((prune4_kernel<T, C, false>), grid, dim3(64 * C), 0, c->stream, c->run_ops, c->run_prog,
                           tips, clv, c->d_scl, c->d_P, c->d_roots, sw, site_lnl, c->d_partial, c->m, c->m_pad,
                           c->n_taxa, (int)grid.x, -1, 0, 0, 0, 1, nullptr, nullptr);
    }
}

template <typename T>
int dispatch_prune4(phylo_ctx* c, dim3 grid, double* site_lnl) {
    switch (c->C) {
    case 1: launch_prune4<T, 1>(c, grid, site_lnl); break;
    case 2: launch_prune4<T, 2>(c, grid, site_lnl); break;
    case 4: launch_prune4<T, 4>(c, grid, site_lnl); break;
    case 8: launch_prune4<T, 8>(c, grid, site_lnl); break;
    default: return fail(c, PHYLO_ERR_UNSUPPORTED, "number of rate categories must be 1, 2, 4 or 8");
    }
    return PHYLO_OK;
}


template <typename T, int C, int V>
void launch_mfma(phylo_ctx* c, dim3 grid, double* site_lnl) {
    const auto* tips = reinterpret_cast<const uint32_t*>(c->d_tips);
    T* clv = reinterpret_cast<T*>(c->d_clv);
    const double* sw = c->have_site_w ? c->d_sitew : nullptr;
    const uint32_t full = c->S == 32 ? 0xFFFFFFFFu : ((1u << c->S) - 1u);
    if (c->run_whole_tree && c->n_taxa <= kMfmaTipLdsTaxa) {
        const int nblk = (int)grid.x;
        const int main_b = split_main_blocks(c, nblk), tail = nblk - main_b;
        if (c->tail_split && c->split_K > 0 && grid.y == 1 && tail > 0) {
            // the main blocks and the tail's pieces, then the tail's top program
            const int base = c->split_base[c->run_full ? 0 : 1], K = c->split_K;
            hipLaunchKernelGGL((prune_mfma_kernel<T, C, V, kMfmaGroups, true>), dim3(main_b + K * tail), dim3(64 * C),
                               0, c->stream, c->run_ops, c->d_prog, tips, clv, c->d_scl, c->d_Pm, c->d_PTm,
                               c->d_roots, sw, site_lnl, c->d_partial, c->m, c->m_pad, c->S, full, c->n_taxa, main_b,
                               c->run_full ? 0 : 1, 0, base, main_b, tail, nullptr, nullptr);
            // (the top program reads a few tips: no LDS staging)
            unsigned* fold = 64 * C == kSumThreads ? c->d_fold : nullptr;
            hipLaunchKernelGGL((prune_mfma_kernel<T, C, V, kMfmaGroups, false>), dim3(tail), dim3(64 * C), 0,
                               c->stream, c->run_ops, c->d_prog, tips, clv, c->d_scl, c->d_Pm, c->d_PTm, c->d_roots,
                               sw, site_lnl, c->d_partial, c->m, c->m_pad, c->S, full, c->n_taxa, tail, base + K,
                               main_b, 0, 0, 1, fold, c->h_total_dev);
            c->folded = fold != nullptr;
        } else {
            hipLaunchKernelGGL((prune_mfma_kernel<T, C, V, kMfmaGroups, true>), grid, dim3(64 * C), 0, c->stream,
                               c->run_ops, c->run_prog, tips, clv, c->d_scl, c->d_Pm, c->d_PTm, c->d_roots, sw,
                               site_lnl, c->d_partial, c->m, c->m_pad, c->S, full, c->n_taxa, (int)grid.x, -1, 0, 0, 0,
                               1, nullptr, nullptr);
        }
    } else if (grid.y > 1) {
        // many programs (an NNI neighbourhood): program-major, grouped by XCD,
        // as in launch_prune4, but one program a workgroup: these ops are
        // long enough that fewer, longer workgroups do not pay (one, two,
        // four programs: 2.976, 2.997, 3.014 ms for a 122-variant
        // neighbourhood of 64 taxa x 20,000 sites)
        const int nblk = (int)grid.x, n_prog = (int)grid.y, nb8 = (nblk + 7) / 8 * 8;
        const int ppw = 1, n_pg = n_prog;
        hipLaunchKernelGGL((prune_mfma_nni_kernel<T, C, V, kMfmaGroups>), dim3(nb8 * n_pg), dim3(64 * C), 0,
                           c->stream, c->run_ops, c->run_prog, tips, clv, c->d_scl, c->d_Pm, c->d_PTm, c->d_roots, sw,
                           site_lnl, c->d_partial, c->m, c->m_pad, c->S, full, c->n_taxa, nblk, -2, ppw, 0, 0, n_prog,
                           nullptr, nullptr);
    } else {
        hipLaunchKernelGGL((prune_mfma_kernel<T, C, V, kMfmaGroups, false>), grid, dim3(64 * C), 0, c->stream,
                           c->run_ops, c->run_prog, tips, clv, c->d_scl, c->d_Pm, c->d_PTm, c->d_roots, sw, site_lnl,
                           c->d_partial, c->m, c->m_pad, c->S, full, c->n_taxa, (int)grid.x, -1, 0, 0, 0, 1, nullptr,
                           nullptr);
    }
}

template <typename T, int C>
int dispatch_mfma_v(phylo_ctx* c, dim3 grid, double* site_lnl) {
    switch (mfma_slots(c->S)) {
    case 2: launch_mfma<T, C, 2>(c, grid, site_lnl); break;
    case 3: launch_mfma<T, C, 3>(c, grid, site_lnl); break;
    case 4: launch_mfma<T, C, 4>(c, grid, site_lnl); break;
    case 5: launch_mfma<T, C, 5>(c, grid, site_lnl); break;
    case 6: launch_mfma<T, C, 6>(c, grid, site_lnl); break;
    case 7: launch_mfma<T, C, 7>(c, grid, site_lnl); break;
    default: launch_mfma<T, C, 8>(c, grid, site_lnl); break;
    }
    return PHYLO_OK;
}

template <typename T>
int dispatch_mfma(phylo_ctx* c, dim3 grid, double* site_lnl) {
    switch (c->C) {
    case 1: return dispatch_mfma_v<T, 1>(c, grid, site_lnl);
    case 2: return dispatch_mfma_v<T, 2>(c, grid, site_lnl);
    case 4: return dispatch_mfma_v<T, 4>(c, grid, site_lnl);
    case 8: return dispatch_mfma_v<T, 8>(c, grid, site_lnl);
    default: return fail(c, PHYLO_ERR_UNSUPPORTED, "number of rate categories must be 1, 2, 4 or 8");
    }
}

// Launches the program set currently in c->ops / c->prog_begin and returns
// one total per program in totals (n_prog values); site_lnl [n_prog][m] is optional.
int run_programs(phylo_ctx* c, const Op* d_ops, const int* d_prog, int n_prog, double* totals,
                 double* site_lnl_host) {
    c->run_ops = d_ops;
    c->run_prog = d_prog;
    const bool mfma = uses_mfma(c);
    const int nblk = c->m_pad / (mfma ? 16 * kMfmaGroups : 64);
    HIPCHK(c, ensure(c->d_partial, c->partial_bytes, (size_t)n_prog * nblk * sizeof(double)));
    if (site_lnl_host) HIPCHK(c, ensure(c->d_sitelnl, c->sitelnl_bytes, (size_t)n_prog * c->m * sizeof(double)));
    if (c->h_total_bytes < (size_t)n_prog * sizeof(double)) {
        if (c->h_total) HIPCHK(c, hipHostFree(c->h_total));
        c->h_total = nullptr;
        HIPCHK(c, hipHostMalloc(reinterpret_cast<void**>(&c->h_total), (size_t)n_prog * sizeof(double),
                                hipHostMallocMapped));
        HIPCHK(c, hipHostGetDevicePointer(reinterpret_cast<void**>(&c->h_total_dev), c->h_total, 0));
        c->h_total_bytes = (size_t)n_prog * sizeof(double);
    }
    dim3 grid(nblk, n_prog);
    double* dsite = site_lnl_host ? c->d_sitelnl : nullptr;
    if (!c->d_fold) {
        // cleared on the context's own stream: a null-stream hipMemset is not
        // ordered with this non-blocking stream, and with other host threads
        // busy on the device the first launch could find the counter unset
        HIPCHK(c, hipMalloc(&c->d_fold, sizeof(unsigned)));
        HIPCHK(c, hipMemsetAsync(c->d_fold, 0, sizeof(unsigned), c->stream));
    }
    c->folded = false;
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    int rc;
    if (c->S <= 4)
        rc = c->prec == PHYLO_PREC_FP64 ? dispatch_prune4<double>(c, grid, dsite) : dispatch_prune4<float>(c, grid, dsite);
    else
        rc = c->prec == PHYLO_PREC_FP64 ? dispatch_mfma<double>(c, grid, dsite) : dispatch_mfma<float>(c, grid, dsite);
    if (rc) return rc;
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    // the totals go straight into mapped pinned memory: no copy (a tail
    // split's top launch has written the total already)
    if (!c->folded) {
        hipLaunchKernelGGL(final_sum_kernel, dim3(n_prog), dim3(kSumThreads), 0, c->stream, c->d_partial, nblk,
                           c->h_total_dev);
        HIPCHK(c, hipGetLastError());
    }
    if (site_lnl_host)
        HIPCHK(c, hipMemcpyAsync(site_lnl_host, c->d_sitelnl, (size_t)n_prog * c->m * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    std::memcpy(totals, c->h_total, (size_t)n_prog * sizeof(double));
    return PHYLO_OK;
}

// ---------------------------------------------------------------------------
// NNI neighbourhoods (tree.cpp:2468-2548, treebender.cpp:897-915)
// ---------------------------------------------------------------------------

template <class TV>
std::vector<int> parents_of(const TV* c) {
    std::vector<int> par(c->n_nodes, -1);
    for (int v = 0; v < c->n_nodes; ++v) {
        if (c->left[v] >= 0) par[c->left[v]] = v;
        if (c->right[v] >= 0) par[c->right[v]] = v;
    }
    return par;
}

template <class TV>
int count_tips(const TV* c, int v) {
    if (v < 0) return 0;
    if (c->left[v] < 0 && c->right[v] < 0) return 1;
    return count_tips(c, c->left[v]) + count_tips(c, c->right[v]);
}

// Variants as (node, side): mode 0 follows treebender --nni all (internal
// branches 2..n-3 numbered in preorder by get_internal_branch_by_number_unrooted,
// tree.cpp:2455-2478; left then right swap); mode 1 is every internal edge of
// the unrooted tree, both root children counting as one edge.
template <class TV>
void enumerate_nni(const TV* c, int mode, std::vector<int>& nodes, std::vector<int>& sides) {
    nodes.clear();
    sides.clear();
    const int r = c->root;
    const int a = c->left[r], b = c->right[r];
    const auto internal = [&](int v) { return v >= 0 && c->left[v] >= 0; };
    if (mode == 0) {
        const int ntips = count_tips(c, r);
        int start = r, offset = 1;
        if (count_tips(c, a) < 2) { start = b; offset = 0; }
        else if (count_tips(c, b) < 2) { start = a; offset = 0; }
        std::vector<int> pre, stack;
        if (start >= 0) stack.push_back(start);
        while (!stack.empty()) {
            int v = stack.back();
            stack.pop_back();
            if (!internal(v)) continue;
            pre.push_back(v);
            if (c->right[v] >= 0) stack.push_back(c->right[v]);
            stack.push_back(c->left[v]);
        }
        for (int br = 2; br <= ntips - 3; ++br) {
            const int idx = br - 1 + offset;
            const int v = idx < (int)pre.size() ? pre[idx] : -1;
            for (int side = 0; side < 2; ++side) { nodes.push_back(v); sides.push_back(side); }
        }
    } else {
        const std::vector<int> par = parents_of(c);
        for (int v = 0; v < c->n_nodes; ++v) {
            if (v == r || !internal(v)) continue;
            if (par[v] == r && (!(internal(a) && internal(b)) || v == b)) continue;
            for (int side = 0; side < 2; ++side) { nodes.push_back(v); sides.push_back(side); }
        }
    }
}

// Appends the partial-recompute program of tree::interchange(v, side == 0)
// (tree.cpp:2480-2518): the two nodes whose children change, then the path to
// the root, each child coming from a stored partial, a tip or a register
// carry.  Returns false when the interchange is a no-op (no swap partner).
template <class TV>
bool append_nni_program(const TV* c, const std::vector<int>& par, int v, int side, std::vector<Op>& ops) {
    if (v < 0 || v == c->root || c->left[v] < 0) return false;
    const int p = par[v], r = c->root;
    const auto mem = [&](int x, int& kind, int& idx) {
        if (x < c->n_taxa) { kind = SRC_TIP; idx = x; }
        else { kind = SRC_CLV; idx = x - c->n_taxa; }
    };
    const int sib = c->left[p] == v ? c->right[p] : c->left[p];
    int partner = -1;
    if (p != r) partner = sib;
    else if (sib >= 0 && c->left[sib] >= 0) partner = c->left[sib];
    if (partner < 0) return false;
    const int tmp = side == 0 ? c->left[v] : c->right[v];
    // v with `partner` in place of `tmp`
    Op o1{};
    const int vl = side == 0 ? partner : c->left[v], vr = side == 0 ? c->right[v] : partner;
    o1.l_node = vl; o1.r_node = vr;
    mem(vl, o1.l_kind, o1.l_idx);
    mem(vr, o1.r_kind, o1.r_idx);
    o1.out_slot = -1;
    o1.flags = 0;
    ops.push_back(o1);
    const int pp = p != r ? p : sib;            // partner's old parent
    // pp with `tmp` in place of `partner` (and v carried when pp == p)
    Op o2{};
    const int ql = c->left[pp] == partner ? tmp : c->left[pp];
    const int qr = c->right[pp] == partner ? tmp : c->right[pp];
    o2.l_node = ql; o2.r_node = qr;
    if (ql == v) { o2.l_kind = SRC_CARRY_A; o2.l_idx = 0; } else mem(ql, o2.l_kind, o2.l_idx);
    if (qr == v) { o2.r_kind = SRC_CARRY_A; o2.r_idx = 0; } else mem(qr, o2.r_kind, o2.r_idx);
    o2.out_slot = -1;
    if (p != r) {
        o2.flags = 0;
        ops.push_back(o2);
        int cur = p;
        while (cur != r) {
            const int up = par[cur];
            Op o{};
            o.l_node = c->left[up];
            o.r_node = c->right[up];
            if (c->left[up] == cur) { o.l_kind = SRC_CARRY_A; o.l_idx = 0; mem(c->right[up], o.r_kind, o.r_idx); }
            else { o.r_kind = SRC_CARRY_A; o.r_idx = 0; mem(c->left[up], o.l_kind, o.l_idx); }
            o.out_slot = -1;
            o.flags = up == r ? OP_ROOT : 0;
            ops.push_back(o);
            cur = up;
        }
    } else {
        // at the root: v waits in B (o2 does not read it), its sibling comes
        // in A, then the root
        ops[ops.size() - 1].flags |= OP_TO_B;
        o2.flags = 0;
        ops.push_back(o2);
        Op o3{};
        o3.l_node = c->left[r];
        o3.r_node = c->right[r];
        o3.l_kind = c->left[r] == v ? SRC_CARRY_B : SRC_CARRY_A;
        o3.r_kind = c->right[r] == v ? SRC_CARRY_B : SRC_CARRY_A;
        o3.l_idx = o3.r_idx = 0;
        o3.out_slot = -1;
        o3.flags = OP_ROOT;
        ops.push_back(o3);
    }
    return true;
}

// Checks a rooted topology over n_taxa tips (nodes 0..n_taxa-1, no children)
// and internal nodes n_taxa..n_nodes-1: one or two valid children each, every
// node but the root with exactly one parent.  Returns the reason, or nullptr.
const char* check_topology(int n_taxa, int n_nodes, int root, const int32_t* left, const int32_t* right) {
    if (n_nodes <= n_taxa || root < n_taxa || root >= n_nodes || !left || !right) return "bad node count or root";
    std::vector<int> indeg(n_nodes, 0);
    for (int v = 0; v < n_nodes; ++v) {
        const bool tip = v < n_taxa;
        if (tip && (left[v] != -1 || right[v] != -1)) return "tips must have no children";
        if (!tip) {
            if (left[v] < 0 || left[v] >= n_nodes || right[v] < -1 || right[v] >= n_nodes || left[v] == right[v] ||
                left[v] == v || right[v] == v)
                return "bad child index";
            ++indeg[left[v]];
            if (right[v] >= 0) ++indeg[right[v]];
        }
    }
    for (int v = 0; v < n_nodes; ++v)
        if ((v == root && indeg[v] != 0) || (v != root && indeg[v] != 1))
            return "every non-root node needs exactly one parent";
    return nullptr;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

extern "C" {

int phylo_version(void) { return 1; }

int phylo_create(int device, phylo_ctx** out) {
    if (!out) return PHYLO_ERR_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return PHYLO_ERR_HIP;
    if (device < 0 || device >= n) return PHYLO_ERR_ARG;
    phylo_ctx* c = new phylo_ctx();
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
        hipDeviceGetAttribute(&c->n_cu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) {
        phylo_destroy(c);
        return PHYLO_ERR_HIP;
    }
    *out = c;
    return PHYLO_OK;
}

void phylo_destroy(phylo_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    void* bufs[] = {c->d_tips, c->d_Pm, c->d_PTm, c->d_clv, c->d_scl, c->d_P, c->d_ops, c->d_prog, c->d_nni_ops, c->d_nni_prog,
                    c->d_roots, c->d_sitew,
                    c->d_sitelnl, c->d_partial, c->d_eig, c->d_fold};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    if (c->h_total) (void)hipHostFree(c->h_total);
    stage_free(c->stage_P);
    stage_free(c->stage_roots);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char* phylo_last_error(const phylo_ctx* c) { return c ? c->err.c_str() : "null context"; }

int phylo_set_precision(phylo_ctx* c, int precision) {
    if (!c) return PHYLO_ERR_ARG;
    if (precision != PHYLO_PREC_FP32 && precision != PHYLO_PREC_FP64) return fail(c, PHYLO_ERR_ARG, "precision");
    if (precision != c->prec) {
        c->prec = precision;
        c->clv_valid = false;
        if (c->d_clv) { (void)hipFree(c->d_clv); c->d_clv = nullptr; c->clv_bytes = 0; }
    }
    return PHYLO_OK;
}

int phylo_set_tips(phylo_ctx* c, int n_taxa, int n_sites, int n_states, const uint32_t* masks) {
    if (!c) return PHYLO_ERR_ARG;
    if (n_taxa < 2 || n_sites < 1 || n_states < 1 || n_states > kMaxStates || !masks)
        return fail(c, PHYLO_ERR_ARG, "set_tips: need n_taxa >= 2, n_sites >= 1, 1 <= n_states <= 32");
    const uint32_t valid = n_states == 32 ? 0xFFFFFFFFu : ((1u << n_states) - 1u);
    for (size_t i = 0; i < (size_t)n_taxa * n_sites; ++i)
        if ((masks[i] & valid) == 0u) {
            // simpleML::check_nodes (simpleML.cpp:52-71): a tip without any state
            char buf[160];
            snprintf(buf, sizeof buf, "taxon %zu lacks a character state for character %zu", i / n_sites, i % n_sites);
            return fail(c, PHYLO_ERR_DATA, buf);
        }
    if (c->have_topo && n_taxa != c->n_taxa) {
        // node numbering (tips 0..n_taxa-1) changed: the topology and the
        // lengths set for the old numbering no longer apply
        c->have_topo = false;
        c->have_eff = false;
    }
    c->n_taxa = n_taxa;
    c->m = n_sites;
    c->S = n_states;
    c->m_pad = (n_sites + kBlock - 1) / kBlock * kBlock;
    c->masks.assign(masks, masks + (size_t)n_taxa * n_sites);
    for (auto& x : c->masks) x &= valid;
    c->have_tips = true;
    c->dirty_tips = true;
    c->clv_valid = false;
    c->dirty_roots = true;
    c->dirty_topo = true;     // slot numbering depends on n_taxa
    c->have_site_w = false;
    c->dirty_site_w = false;
    return PHYLO_OK;
}

int phylo_set_site_weights(phylo_ctx* c, const double* w) {
    if (!c) return PHYLO_ERR_ARG;
    if (!c->have_tips) return fail(c, PHYLO_ERR_STATE, "set_site_weights before set_tips");
    // the stored partials do not depend on the weights, but the cached total
    // of the last evaluation (the NNI base and no-op variants) does
    c->clv_valid = false;
    if (!w) { c->have_site_w = false; c->dirty_site_w = false; return PHYLO_OK; }
    for (int s = 0; s < c->m; ++s)
        if (!std::isfinite(w[s])) return fail(c, PHYLO_ERR_ARG, "set_site_weights: non-finite weight");
    c->site_w.assign(w, w + c->m);
    c->have_site_w = true;
    c->dirty_site_w = true;
    return PHYLO_OK;
}

int phylo_set_topology(phylo_ctx* c, int n_nodes, int root, const int32_t* left, const int32_t* right) {
    if (!c) return PHYLO_ERR_ARG;
    if (!c->have_tips) return fail(c, PHYLO_ERR_STATE, "set_topology before set_tips");
    const char* why = check_topology(c->n_taxa, n_nodes, root, left, right);
    if (why) return fail(c, PHYLO_ERR_ARG, std::string("set_topology: ") + why);
    c->n_nodes = n_nodes;
    c->root = root;
    c->left.assign(left, left + n_nodes);
    c->right.assign(right, right + n_nodes);
    c->have_topo = true;
    c->dirty_topo = true;
    c->clv_valid = false;
    c->have_eff = false;
    c->dirty_P = true;
    return PHYLO_OK;
}

int phylo_set_edge_lengths(phylo_ctx* c, const double* eff) {
    if (!c) return PHYLO_ERR_ARG;
    if (!c->have_topo) return fail(c, PHYLO_ERR_STATE, "set_edge_lengths before set_topology");
    if (!eff) return fail(c, PHYLO_ERR_ARG, "set_edge_lengths: null");
    for (int v = 0; v < c->n_nodes; ++v)
        if (v != c->root && !std::isfinite(eff[v])) return fail(c, PHYLO_ERR_ARG, "set_edge_lengths: non-finite length");
    c->eff.assign(eff, eff + c->n_nodes);
    c->eff[c->root] = 0.0;
    c->have_eff = true;
    c->dirty_P = true;
    c->clv_valid = false;
    return PHYLO_OK;
}

int phylo_set_model(phylo_ctx* c, int n_states, const double* Q, const double* root_w, const double* pi,
                    int p_builder) {
    if (!c) return PHYLO_ERR_ARG;
    if (n_states < 1 || n_states > kMaxStates || !Q || !root_w) return fail(c, PHYLO_ERR_ARG, "set_model: bad arguments");
    if (p_builder != PHYLO_P_REFCOMPAT && p_builder != PHYLO_P_EIGEN && p_builder != PHYLO_P_PADE)
        return fail(c, PHYLO_ERR_ARG, "set_model: unknown P builder");
    for (int i = 0; i < n_states * n_states; ++i)
        if (!std::isfinite(Q[i])) return fail(c, PHYLO_ERR_ARG, "set_model: non-finite Q");
    c->Q.assign(Q, Q + (size_t)n_states * n_states);
    c->root_w.assign(root_w, root_w + n_states);
    c->S_model = n_states;
    c->builder = p_builder;
    if (p_builder == PHYLO_P_EIGEN) {
        if (!pi) return fail(c, PHYLO_ERR_ARG, "set_model: the eigen builder needs pi");
        c->pi.assign(pi, pi + n_states);
        if (!phylo_math::is_reversible(n_states, Q, pi))
            return fail(c, PHYLO_ERR_ARG, "set_model: Q is not reversible with respect to pi; use PHYLO_P_PADE");
        if (!phylo_math::reversible_eigen(n_states, Q, pi, c->V, c->lam, c->Vinv))
            return fail(c, PHYLO_ERR_ARG, "set_model: eigen-decomposition failed (pi must be positive)");
    }
    c->have_model = true;
    c->dirty_P = true;
    c->clv_valid = false;
    c->dirty_roots = true;
    return PHYLO_OK;
}

int phylo_set_categories(phylo_ctx* c, int n_cat, const double* rates, const double* weights) {
    if (!c) return PHYLO_ERR_ARG;
    if (n_cat < 1 || n_cat > kMaxCat || !rates || !weights) return fail(c, PHYLO_ERR_ARG, "set_categories: 1..8 categories");
    for (int k = 0; k < n_cat; ++k)
        if (!(rates[k] >= 0.0) || !(weights[k] >= 0.0) || !std::isfinite(rates[k]))
            return fail(c, PHYLO_ERR_ARG, "set_categories: rates and weights must be finite and >= 0");
    c->C = n_cat;
    c->cat_rate.assign(rates, rates + n_cat);
    c->cat_w.assign(weights, weights + n_cat);
    c->dirty_P = true;
    c->clv_valid = false;
    c->dirty_roots = true;
    return PHYLO_OK;
}

namespace {
// One evaluation of the current inputs; `full` stores every internal node's
// partial, otherwise only those the walk reads back.
int evaluate(phylo_ctx* c, bool full, double* site_lnl, double* total) {
    int rc = prepare(c);
    if (rc) return rc;
    c->clv_valid = false;
    c->run_whole_tree = true;
    c->run_full = full;
    rc = run_programs(c, c->d_ops, c->d_prog + (full ? 0 : 1), 1, total, site_lnl);
    c->run_whole_tree = false;
    if (rc) return rc;
    c->clv_valid = full;
    c->base_total = *total;
    return PHYLO_OK;
}
}  // namespace

int phylo_set_tail_split(phylo_ctx* c, int enable) {
    if (!c) return PHYLO_ERR_ARG;
    c->tail_split = enable != 0;
    c->clv_valid = false;      // the tail's stored partials depend on the walk used
    return PHYLO_OK;
}

int phylo_set_keep_partials(phylo_ctx* c, int keep) {
    if (!c) return PHYLO_ERR_ARG;
    c->keep_partials = keep != 0;
    return PHYLO_OK;
}

int phylo_loglik(phylo_ctx* c, double* site_lnl, double* total) {
    if (!c) return PHYLO_ERR_ARG;
    if (!total) return fail(c, PHYLO_ERR_ARG, "loglik: total is required");
    return evaluate(c, c->keep_partials, site_lnl, total);
}

int phylo_node_partials(phylo_ctx* c, int node, double* clv, int32_t* scale) {
    if (!c) return PHYLO_ERR_ARG;
    if (!c->d_clv || c->ops.empty()) return fail(c, PHYLO_ERR_STATE, "node_partials before loglik");
    if (node < c->n_taxa || node >= c->n_nodes) return fail(c, PHYLO_ERR_ARG, "node_partials: not an internal node");
    if (!c->clv_valid) {
        // the last evaluation stored only what its walk read back (or the
        // inputs changed since): one evaluation that stores every partial
        double t = 0.0;
        int rc = evaluate(c, true, nullptr, &t);
        if (rc) return rc;
    }
    const int slot = node - c->n_taxa;
    const int Sp = clv_states(c);
    const size_t n = (size_t)c->C * c->m_pad * Sp;
    HIPCHK(c, hipSetDevice(c->device));
    std::vector<double> buf(n);
    if (c->prec == PHYLO_PREC_FP64) {
        HIPCHK(c, hipMemcpyAsync(buf.data(), reinterpret_cast<double*>(c->d_clv) + slot * n, n * sizeof(double),
                                 hipMemcpyDeviceToHost, c->stream));
    } else {
        std::vector<float> fbuf(n);
        HIPCHK(c, hipMemcpyAsync(fbuf.data(), reinterpret_cast<float*>(c->d_clv) + slot * n, n * sizeof(float),
                                 hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        std::copy(fbuf.begin(), fbuf.end(), buf.begin());
    }
    // one scaler per (category, site): returned at the site's largest, the
    // other categories' values shifted to it
    std::vector<int> csc((size_t)c->C * c->m_pad);
    HIPCHK(c, hipMemcpyAsync(csc.data(), c->d_scl + (size_t)slot * c->C * c->m_pad, csc.size() * sizeof(int),
                             hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const bool mfma = uses_mfma(c);
    const int V = mfma_slots(c->S), n_sg = c->m_pad / 16;
    // DNA: [cat][site][4]; wide: [cat][site group][V][64 lanes], state i of
    // site s in slot i / 4 of lane (i % 4) * 16 + s % 16
    const auto at = [&](int k, int s, int i) -> double {
        return mfma ? buf[(((size_t)k * n_sg + s / 16) * V + i / 4) * 64 + (i % 4) * 16 + s % 16]
                    : buf[((size_t)k * c->m_pad + s) * Sp + i];
    };
    for (int s = 0; s < c->m; ++s) {
        int M = csc[s];
        for (int k = 1; k < c->C; ++k) M = std::max(M, csc[(size_t)k * c->m_pad + s]);
        if (scale) scale[s] = M;
        for (int k = 0; k < c->C; ++k) {
            const int d = csc[(size_t)k * c->m_pad + s] - M;
            for (int i = 0; i < c->S; ++i) clv[((size_t)s * c->C + k) * c->S + i] = std::ldexp(at(k, s, i), d);
        }
    }
    return PHYLO_OK;
}

int phylo_get_pmatrix(phylo_ctx* c, int node, double* P) {
    if (!c || !P) return PHYLO_ERR_ARG;
    if (node < 0 || node >= c->n_nodes) return fail(c, PHYLO_ERR_ARG, "get_pmatrix: node out of range");
    if (c->dirty_P || !c->d_P) {
        int rc = prepare(c);
        if (rc) return rc;
    }
    const int Sp = padded_states(c->S_model), S = c->S_model;
    std::vector<double> buf((size_t)c->C * Sp * Sp);
    // the P build runs on the context's non-blocking stream: copy on that
    // stream, after it, not on the null stream
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipMemcpyAsync(buf.data(), c->d_P + (size_t)node * c->C * Sp * Sp, buf.size() * sizeof(double),
                             hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int k = 0; k < c->C; ++k)
        for (int i = 0; i < S; ++i)
            for (int j = 0; j < S; ++j) P[((size_t)k * S + i) * S + j] = buf[((size_t)k * Sp + i) * Sp + j];
    return PHYLO_OK;
}

int phylo_nni_score(phylo_ctx* c, int mode, int max_variants, double* lnl, int32_t* node, int32_t* side,
                    int* n_variants, double* base_lnl) {
    if (!c || !n_variants) return PHYLO_ERR_ARG;
    if (mode != 0 && mode != 1) return fail(c, PHYLO_ERR_ARG, "nni_score: mode must be 0 or 1");
    int rc = prepare(c);
    if (rc) return rc;
    double base = 0.0;
    if (!c->clv_valid) {
        rc = evaluate(c, true, nullptr, &base);   // the variants read every stored partial
        if (rc) return rc;
    }
    base = c->base_total;
    std::vector<int> nodes, sides;
    enumerate_nni(c, mode, nodes, sides);
    const int nv = (int)nodes.size();
    *n_variants = nv;
    if (base_lnl) *base_lnl = base;
    if (nv == 0 || !lnl) return PHYLO_OK;   // count query
    if (max_variants < nv) return fail(c, PHYLO_ERR_ARG, "nni_score: output arrays too small");
    const std::vector<int> par = parents_of(c);
    std::vector<Op> ops;
    std::vector<int> begin{0}, which;
    for (int k = 0; k < nv; ++k) {
        if (append_nni_program(c, par, nodes[k], sides[k], ops)) {
            begin.push_back((int)ops.size());
            which.push_back(k);
        }
        lnl[k] = base;   // no-op interchanges keep the base tree
        if (node) node[k] = nodes[k];
        if (side) side[k] = sides[k];
    }
    const int n_prog = (int)which.size();
    if (n_prog == 0) return PHYLO_OK;
    if (!root_ops_last(ops, begin)) return fail(c, PHYLO_ERR_STATE, "internal: root op not last");
    HIPCHK(c, upload_ops(c->d_nni_ops, c->nni_ops_bytes, ops, c->stream));
    HIPCHK(c, ensure(c->d_nni_prog, c->nni_prog_bytes, begin.size() * sizeof(int)));
    HIPCHK(c, hipMemcpyAsync(c->d_nni_prog, begin.data(), begin.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
    std::vector<double> tot(n_prog);
    rc = run_programs(c, c->d_nni_ops, c->d_nni_prog, n_prog, tot.data(), nullptr);
    if (rc) return rc;
    for (int q = 0; q < n_prog; ++q) lnl[which[q]] = tot[q];
    c->nni_ops_last = (long)ops.size();
    return PHYLO_OK;
}

int phylo_simulate(phylo_ctx* c, int n_sites, const double* root_freq, uint64_t seed, uint8_t* states) {
    if (!c) return PHYLO_ERR_ARG;
    if (n_sites < 1 || !root_freq || !states) return fail(c, PHYLO_ERR_ARG, "simulate: bad arguments");
    if (!c->have_topo) return fail(c, PHYLO_ERR_STATE, "simulate: phylo_set_topology has not been called");
    if (!c->have_eff) return fail(c, PHYLO_ERR_STATE, "simulate: phylo_set_edge_lengths has not been called");
    if (!c->have_model) return fail(c, PHYLO_ERR_STATE, "simulate: phylo_set_model has not been called");
    const int S = c->S_model, n = c->n_nodes;
    for (int j = 0; j < S; ++j)
        if (!std::isfinite(root_freq[j])) return fail(c, PHYLO_ERR_ARG, "simulate: non-finite root frequency");
    HIPCHK(c, hipSetDevice(c->device));
    if (c->dirty_P) {
        int rc = upload_pmatrices(c);
        if (rc) return rc;
    }
    // preorder and parents on the host (every node reachable: checked by set_topology)
    std::vector<int> pre, parent(n, -1), stack{c->root};
    pre.reserve(n);
    while (!stack.empty()) {
        const int v = stack.back();
        stack.pop_back();
        pre.push_back(v);
        if (c->right[v] >= 0) { parent[c->right[v]] = v; stack.push_back(c->right[v]); }
        if (c->left[v] >= 0) { parent[c->left[v]] = v; stack.push_back(c->left[v]); }
    }
    if ((int)pre.size() != n) return fail(c, PHYLO_ERR_ARG, "simulate: nodes unreachable from the root");
    double wsum = 0.0;
    for (int k = 0; k < c->C; ++k) wsum += c->cat_w[k];
    std::vector<double> aux(S + c->C);
    for (int j = 0; j < S; ++j) aux[j] = root_freq[j];
    for (int k = 0; k < c->C; ++k) aux[S + k] = wsum > 0 ? c->cat_w[k] / wsum : 1.0 / c->C;
    int* d_idx = nullptr;
    double* d_aux = nullptr;
    uint8_t* d_out = nullptr;
    const size_t n_out = (size_t)n * n_sites;
    auto cleanup = [&]() {
        if (d_idx) (void)hipFree(d_idx);
        if (d_aux) (void)hipFree(d_aux);
        if (d_out) (void)hipFree(d_out);
    };
    std::vector<int> idx(pre);
    idx.insert(idx.end(), parent.begin(), parent.end());
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&d_idx), idx.size() * sizeof(int));
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&d_aux), aux.size() * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&d_out), n_out);
    if (e == hipSuccess) e = hipMemcpyAsync(d_idx, idx.data(), idx.size() * sizeof(int), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d_aux, aux.data(), aux.size() * sizeof(double), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(simulate_kernel, dim3((unsigned)((n_sites + 255) / 256)), dim3(256), 0, c->stream, d_idx,
                           d_idx + n, n, n_sites, S, padded_states(S), c->C, c->d_P, d_aux, d_aux + S, seed, d_out);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(states, d_out, n_out, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    cleanup();
    if (e != hipSuccess) return fail(c, PHYLO_ERR_HIP, std::string("simulate: ") + hipGetErrorString(e));
    return PHYLO_OK;
}

int phylo_nni_ops(phylo_ctx* c, long* n_ops) {
    if (!c || !n_ops) return PHYLO_ERR_ARG;
    *n_ops = c->nni_ops_last;
    return PHYLO_OK;
}

int phylo_synchronize(phylo_ctx* c) {
    if (!c) return PHYLO_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipDeviceSynchronize());
    return PHYLO_OK;
}

int phylo_last_kernel_ms(phylo_ctx* c, double* ms) {
    if (!c || !ms) return PHYLO_ERR_ARG;
    *ms = c->last_ms;
    return PHYLO_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Allocation-free evaluation (phylo_loglik_noalloc)
// ---------------------------------------------------------------------------

namespace {
constexpr size_t kScratchBytes = size_t(64) << 20;
constexpr int kScratchDevices = 64;
}  // namespace

// The whole device workspace of phylo_loglik_noalloc: inputs staged from the
// host, P matrices, stored partials and scalers, block sums and outputs.  One
// copy per GPU, held until the process ends.  On the host this name is a
// stand-in, not a device address: host code uses the per-device address
// from hipGetSymbolAddress (scratch_base) and never the name itself.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

std::mutex g_scratch_mu;
unsigned char* g_scratch_base[kScratchDevices] = {};

// Device address of `scratch` on `dev`, looked up once per device with that
// device current (each GPU holds its own copy at its own address).
hipError_t scratch_base(int dev, unsigned char** out) {
    if (dev < 0 || dev >= kScratchDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_base[dev]) {
        hipError_t e = hipSetDevice(dev);
        if (e != hipSuccess) return e;
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_base[dev] = static_cast<unsigned char*>(p);
    }
    *out = g_scratch_base[dev];
    return hipSuccess;
}

// The topology of one call, in the shape build_full_ops reads.
struct TopoView {
    int n_taxa = 0, n_nodes = 0, root = -1;
    std::vector<int> left, right;
};

template <int C>
void launch_noalloc(hipStream_t st, int nblk, bool tip_lds, const Op* ops, const int* prog, const uint8_t* tips,
                    float* clv, int* scl, const double* P, const double* roots, double* site, double* partial, int m,
                    int m_pad, int n_taxa) {
    if (tip_lds)
        hipLaunchKernelGGL((prune4_kernel<float, C, true>), dim3(nblk), dim3(64 * C), tip_lds_bytes(n_taxa), st, ops,
                           prog, tips, clv, scl,
                           P, roots, nullptr, site, partial, m, m_pad, n_taxa, nblk, -1, 0, 0, 0, 1, nullptr, nullptr);
    else
        hipLaunchKernelGGL((prune4_kernel<float, C, false>), dim3(nblk), dim3(64 * C), 0, st, ops, prog, tips, clv,
                           scl, P, roots, nullptr, site, partial, m, m_pad, n_taxa, nblk, -1, 0, 0, 0, 1, nullptr,
                           nullptr);
}

}  // namespace

extern "C" int phylo_loglik_noalloc(phylo_ctx* c, int n_taxa, int n_sites, int n_states, const uint32_t* masks,
                                    int n_nodes, int root, const int32_t* left, const int32_t* right,
                                    const double* eff_len, const double* Q, const double* root_weights,
                                    const double* pi, int n_cat, const double* cat_rates, const double* cat_weights,
                                    double* site_lnl, double* total) {
    if (!c) return PHYLO_ERR_ARG;
    if (!total || !masks || !eff_len || !Q || !root_weights || !pi || !cat_rates || !cat_weights)
        return fail(c, PHYLO_ERR_ARG, "loglik_noalloc: null argument");
    if (n_states < 1 || n_states > 4)
        return fail(c, PHYLO_ERR_UNSUPPORTED, "loglik_noalloc: 1 to 4 states only (the DNA walker)");
    if (n_cat != 1 && n_cat != 2 && n_cat != 4 && n_cat != 8)
        return fail(c, PHYLO_ERR_UNSUPPORTED, "loglik_noalloc: number of rate categories must be 1, 2, 4 or 8");
    if (n_taxa < 2 || n_sites < 1) return fail(c, PHYLO_ERR_ARG, "loglik_noalloc: need n_taxa >= 2, n_sites >= 1");
    const char* why = check_topology(n_taxa, n_nodes, root, left, right);
    if (why) return fail(c, PHYLO_ERR_ARG, std::string("loglik_noalloc: ") + why);
    for (int v = 0; v < n_nodes; ++v)
        if (v != root && !std::isfinite(eff_len[v])) return fail(c, PHYLO_ERR_ARG, "loglik_noalloc: non-finite length");
    for (int i = 0; i < n_states * n_states; ++i)
        if (!std::isfinite(Q[i])) return fail(c, PHYLO_ERR_ARG, "loglik_noalloc: non-finite Q");
    for (int k = 0; k < n_cat; ++k)
        if (!(cat_rates[k] >= 0.0) || !(cat_weights[k] >= 0.0) || !std::isfinite(cat_rates[k]))
            return fail(c, PHYLO_ERR_ARG, "loglik_noalloc: rates and weights must be finite and >= 0");
    const uint32_t valid = (1u << n_states) - 1u;
    for (size_t i = 0; i < (size_t)n_taxa * n_sites; ++i)
        if ((masks[i] & valid) == 0u) {
            char buf[160];
            snprintf(buf, sizeof buf, "taxon %zu lacks a character state for character %zu", i / n_sites,
                     i % n_sites);
            return fail(c, PHYLO_ERR_DATA, buf);
        }
    if (!phylo_math::is_reversible(n_states, Q, pi))
        return fail(c, PHYLO_ERR_ARG, "loglik_noalloc: Q is not reversible with respect to pi");
    std::vector<double> V, lam, Vinv;
    if (!phylo_math::reversible_eigen(n_states, Q, pi, V, lam, Vinv))
        return fail(c, PHYLO_ERR_ARG, "loglik_noalloc: eigen-decomposition failed (pi must be positive)");

    // the lean program of phylo_loglik, its stored partials renumbered densely
    TopoView tv;
    tv.n_taxa = n_taxa;
    tv.n_nodes = n_nodes;
    tv.root = root;
    tv.left.assign(left, left + n_nodes);
    tv.right.assign(right, right + n_nodes);
    std::vector<Op> ops;
    if (!build_full_ops(&tv, ops)) return fail(c, PHYLO_ERR_ARG, "loglik_noalloc: internal nodes unreachable");
    if (!root_ops_last(ops, {0, (int)ops.size()})) return fail(c, PHYLO_ERR_STATE, "internal: root op not last");
    drop_unread_stores(ops, 0, ops.size());
    std::vector<int> dense(n_nodes - n_taxa, -1);
    int n_store = 0;
    for (Op& o : ops)
        if (o.out_slot >= 0) {
            dense[o.out_slot] = n_store;
            o.out_slot = n_store++;
        }
    for (Op& o : ops) {
        if (o.l_kind == SRC_CLV) o.l_idx = dense[o.l_idx];
        if (o.r_kind == SRC_CLV) o.r_idx = dense[o.r_idx];
        if (o.l_idx < 0 || o.r_idx < 0) return fail(c, PHYLO_ERR_ARG, "loglik_noalloc: program error");
    }

    // workspace layout, every region 256-byte aligned
    const int m_pad = (n_sites + kBlock - 1) / kBlock * kBlock, nblk = m_pad / 64, C = n_cat;
    const size_t n_eig = V.size() + Vinv.size() + lam.size() + (size_t)C + (size_t)n_nodes;
    const size_t n_roots = kMaxStates + kMaxCat;
    size_t off = 0;
    const auto take = [&](size_t bytes) {
        const size_t at = off;
        off += (bytes + 255) & ~size_t(255);
        return at;
    };
    const size_t o_tips = take((size_t)n_taxa * m_pad);
    const size_t o_ops = take((ops.size() + 1) * sizeof(Op));       // and a zero op after the program
    const size_t o_prog = take(2 * sizeof(int));
    const size_t o_eig = take(n_eig * sizeof(double));
    const size_t o_roots = take(n_roots * sizeof(double));
    const size_t in_bytes = off;                                    // staged from the host in one copy
    const size_t o_P = take((size_t)n_nodes * C * 16 * sizeof(double));
    const size_t o_clv = take((size_t)n_store * C * m_pad * 4 * sizeof(float));
    const size_t o_scl = take((size_t)n_store * C * m_pad * sizeof(int));
    const size_t o_part = take((size_t)nblk * sizeof(double));
    const size_t o_site = take(site_lnl ? (size_t)n_sites * sizeof(double) : 0);
    const size_t o_total = take(sizeof(double));
    if (off > kScratchBytes) {
        char buf[200];
        snprintf(buf, sizeof buf, "loglik_noalloc: this problem needs %zu bytes of workspace, more than the %zu-byte "
                 "scratch array", off, kScratchBytes);
        return fail(c, PHYLO_ERR_ARG, buf);
    }

    std::vector<unsigned char> in(in_bytes, 0);
    for (int t = 0; t < n_taxa; ++t) {
        uint8_t* row = in.data() + o_tips + (size_t)t * m_pad;
        for (int s = 0; s < n_sites; ++s) row[s] = (uint8_t)(masks[(size_t)t * n_sites + s] & valid);
        std::fill(row + n_sites, row + m_pad, (uint8_t)0x0F);     // padding sites: all states
    }
    order_children(ops);
    std::memcpy(in.data() + o_ops, ops.data(), ops.size() * sizeof(Op));
    const int prog[2] = {0, (int)ops.size()};
    std::memcpy(in.data() + o_prog, prog, sizeof prog);
    double* eig = reinterpret_cast<double*>(in.data() + o_eig);
    double* w = std::copy(V.begin(), V.end(), eig);
    w = std::copy(Vinv.begin(), Vinv.end(), w);
    w = std::copy(lam.begin(), lam.end(), w);
    w = std::copy(cat_rates, cat_rates + C, w);
    std::copy(eff_len, eff_len + n_nodes, w);
    w[root] = 0.0;
    double* roots = reinterpret_cast<double*>(in.data() + o_roots);
    for (int i = 0; i < n_states; ++i) roots[i] = root_weights[i];
    for (int k = 0; k < C; ++k) roots[kRootCatOff + k] = cat_weights[k];

    HIPCHK(c, hipSetDevice(c->device));
    unsigned char* base = nullptr;
    HIPCHK(c, scratch_base(c->device, &base));
    hipStream_t st = c->stream;
    // nothing in the workspace has to start at zero: every stored partial,
    // scaler and block sum is written before it is read
    HIPCHK(c, hipMemcpyAsync(base, in.data(), in_bytes, hipMemcpyHostToDevice, st));
    const double* dV = reinterpret_cast<const double*>(base + o_eig);
    const double* dVi = dV + V.size();
    const double* dl = dVi + Vinv.size();
    const double* dr = dl + lam.size();
    const double* d_eff = dr + C;
    double* dP = reinterpret_cast<double*>(base + o_P);
    const long nP = (long)n_nodes * C * 16;
    hipLaunchKernelGGL(pbuild_eigen_kernel, dim3((unsigned)((nP + 255) / 256)), dim3(256), 0, st, dV, dVi, dl, d_eff,
                       dr, n_states, 4, C, n_nodes, dP);
    HIPCHK(c, hipGetLastError());
    const Op* d_ops = reinterpret_cast<const Op*>(base + o_ops);
    const int* d_prog = reinterpret_cast<const int*>(base + o_prog);
    const uint8_t* d_tips = base + o_tips;
    float* d_clv = reinterpret_cast<float*>(base + o_clv);
    int* d_scl = reinterpret_cast<int*>(base + o_scl);
    const double* d_roots = reinterpret_cast<const double*>(base + o_roots);
    double* d_site = site_lnl ? reinterpret_cast<double*>(base + o_site) : nullptr;
    double* d_part = reinterpret_cast<double*>(base + o_part);
    double* d_total = reinterpret_cast<double*>(base + o_total);
    const bool tip_lds = n_taxa <= kTipLdsTaxa;
    switch (C) {
    case 1: launch_noalloc<1>(st, nblk, tip_lds, d_ops, d_prog, d_tips, d_clv, d_scl, dP, d_roots, d_site, d_part,
                              n_sites, m_pad, n_taxa); break;
    case 2: launch_noalloc<2>(st, nblk, tip_lds, d_ops, d_prog, d_tips, d_clv, d_scl, dP, d_roots, d_site, d_part,
                              n_sites, m_pad, n_taxa); break;
    case 4: launch_noalloc<4>(st, nblk, tip_lds, d_ops, d_prog, d_tips, d_clv, d_scl, dP, d_roots, d_site, d_part,
                              n_sites, m_pad, n_taxa); break;
    default: launch_noalloc<8>(st, nblk, tip_lds, d_ops, d_prog, d_tips, d_clv, d_scl, dP, d_roots, d_site, d_part,
                               n_sites, m_pad, n_taxa); break;
    }
    HIPCHK(c, hipGetLastError());
    hipLaunchKernelGGL(final_sum_kernel, dim3(1), dim3(kSumThreads), 0, st, d_part, nblk, d_total);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(total, d_total, sizeof(double), hipMemcpyDeviceToHost, st));
    if (site_lnl)
        HIPCHK(c, hipMemcpyAsync(site_lnl, d_site, (size_t)n_sites * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipStreamSynchronize(st));
    return PHYLO_OK;
}

// ---------------------------------------------------------------------------
// Batched scoring of many independent problems (+ NNI neighbourhoods)
// ---------------------------------------------------------------------------

struct phylo_batch {
    struct Inst {
        int n_taxa = 0, n_nodes = 0, root = -1, m = 0, m_pad = 0, S = 4, nni_mode = -1;
        std::vector<int> left, right;
        std::vector<uint32_t> masks;
        std::vector<double> P;      // [n_nodes][C][16]
        double roots[kMaxStates + kMaxCat];
        long tips_off = 0, clv_off = 0, scl_off = 0, P_off = 0;
        int roots_off = 0;
        int full_prog = -1;
        std::vector<int> nni_prog, nni_node, nni_side;   // nni_prog[k] = -1: no-op variant
        long full_ops = 0, nni_ops = 0;
    };
    int device = 0;
    int C = 4, prec = PHYLO_PREC_FP32;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string err;
    std::vector<Inst> inst;
    bool finalized = false;
    int cat_waves = 4;                // category tiling: C waves per workgroup, or 1
    int per_cu_cap = 0;               // resident workgroups per CU of the persistent grid (0: all that fit)
    int max_per_cu = 1;               // all that fit, for the current tiling
    int blocks_per_item = 2, grid = 0, n_items1 = 0, n_items2 = 0, n_seg1 = 0, n_seg2 = 0, n_var = 0;
    double updates1 = 0, updates2 = 0, last_ms = 0;
    std::vector<int> var_offsets;     // per instance into the flat variant list
    // device
    uint8_t* d_tips = nullptr;
    void* d_clv = nullptr;
    int* d_scl = nullptr;
    double* d_P = nullptr;
    double* d_roots = nullptr;
    Op* d_ops = nullptr;
    int* d_prog = nullptr;
    BatchInst* d_inst = nullptr;
    WorkItem* d_items1 = nullptr;
    WorkItem* d_items2 = nullptr;
    double* d_partial = nullptr;
    int* d_seg1 = nullptr;
    int* d_seg2 = nullptr;
    double* d_out = nullptr;
    double* h_out = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // host copies used to build work lists
    std::vector<WorkItem> items1, items2;
    std::vector<int> seg1, seg2;
    std::vector<int> var_prog_of;     // flat variant -> program (-1 = no-op)
};

namespace {

int fail(phylo_batch* b, int code, const std::string& msg) {
    if (b) b->err = msg;
    return code;
}

template <typename T, int C>
void launch_batch(phylo_batch* b, const WorkItem* items, int n_items, int bpi) {
    if (b->cat_waves == 1)
        hipLaunchKernelGGL((batch4_kernel<T, C, 1>), dim3(b->grid), dim3(64), 0, b->stream, items, n_items,
                           b->d_inst, b->d_prog, b->d_ops, b->d_tips, reinterpret_cast<T*>(b->d_clv), b->d_scl,
                           b->d_P, b->d_roots, b->d_partial, bpi);
    else
        hipLaunchKernelGGL((batch4_kernel<T, C, C>), dim3(b->grid), dim3(64 * C), 0, b->stream, items, n_items,
                           b->d_inst, b->d_prog, b->d_ops, b->d_tips, reinterpret_cast<T*>(b->d_clv), b->d_scl,
                           b->d_P, b->d_roots, b->d_partial, bpi);
}

// The persistent grid of the current category tiling: as many workgroups as
// can be resident at once.
int batch_grid(phylo_batch* b) {
    hipDeviceProp_t prop;
    HIPCHK(b, hipGetDeviceProperties(&prop, b->device));
    const bool f64 = b->prec == PHYLO_PREC_FP64, one = b->cat_waves == 1;
    const void* kfn = nullptr;
#define PHYLO_BATCH_KFN(CC)                                                                                   \
    kfn = f64 ? (one ? (const void*)batch4_kernel<double, CC, 1> : (const void*)batch4_kernel<double, CC, CC>) \
              : (one ? (const void*)batch4_kernel<float, CC, 1> : (const void*)batch4_kernel<float, CC, CC>)
    switch (b->C) {
    case 1: PHYLO_BATCH_KFN(1); break;
    case 2: PHYLO_BATCH_KFN(2); break;
    case 4: PHYLO_BATCH_KFN(4); break;
    default: PHYLO_BATCH_KFN(8); break;
    }
#undef PHYLO_BATCH_KFN
    int per_cu = 1;
    HIPCHK(b, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kfn, 64 * b->cat_waves, 0));
    b->max_per_cu = std::max(1, per_cu);
    b->grid = std::min(b->max_per_cu, b->per_cu_cap > 0 ? b->per_cu_cap : b->max_per_cu) * prop.multiProcessorCount;
    return PHYLO_OK;
}

int enqueue_batch_phase(phylo_batch* b, int phase, int bpi) {
    const WorkItem* items = phase == 1 ? b->d_items1 : b->d_items2;
    const int n = phase == 1 ? b->n_items1 : b->n_items2;
    if (n == 0) return PHYLO_OK;
    const bool f64 = b->prec == PHYLO_PREC_FP64;
    switch (b->C) {
    case 1: f64 ? launch_batch<double, 1>(b, items, n, bpi) : launch_batch<float, 1>(b, items, n, bpi); break;
    case 2: f64 ? launch_batch<double, 2>(b, items, n, bpi) : launch_batch<float, 2>(b, items, n, bpi); break;
    case 4: f64 ? launch_batch<double, 4>(b, items, n, bpi) : launch_batch<float, 4>(b, items, n, bpi); break;
    default: f64 ? launch_batch<double, 8>(b, items, n, bpi) : launch_batch<float, 8>(b, items, n, bpi); break;
    }
    return PHYLO_OK;
}

// Work lists for a given number of 64-site blocks per item.  Items are
// ordered by decreasing estimated cost (program length x sites) so the
// round-robin assignment to workgroups finishes evenly; partial sums are
// placed by segment.
void build_work_lists(phylo_batch* b, int bpi, const std::vector<int>& prog_len) {
    struct Cand { long cost; WorkItem w; };
    std::vector<Cand> c1, c2;
    b->seg1.assign(1, 0);
    b->seg2.assign(1, 0);
    int out = 0;
    for (size_t i = 0; i < b->inst.size(); ++i) {
        const auto& I = b->inst[i];
        const int nb = I.m_pad / 64;
        for (int blk = 0; blk < nb; blk += bpi) {
            const int sites = std::min(bpi, nb - blk) * 64;
            c1.push_back({(long)prog_len[I.full_prog] * sites, {(int)i, I.full_prog, blk, out++}});
        }
        b->seg1.push_back(out);
    }
    const int base2 = out;
    // a variant's blocks keep consecutive partial slots (its segment), but
    // the list runs block by block over the variants, so that items of equal
    // cost that run side by side are variants of one site block and share
    // its stored partials in L2
    for (size_t i = 0; i < b->inst.size(); ++i) {
        const auto& I = b->inst[i];
        const int nb = I.m_pad / 64, n_items = (nb + bpi - 1) / bpi;
        std::vector<int> first;
        for (int pg : I.nni_prog) {
            if (pg < 0) continue;
            first.push_back(out);
            out += n_items;
            b->seg2.push_back(out - base2);
        }
        for (int blk = 0, q = 0; blk < nb; blk += bpi, ++q) {
            const int sites = std::min(bpi, nb - blk) * 64;
            size_t v = 0;
            for (int pg : I.nni_prog) {
                if (pg < 0) continue;
                c2.push_back({(long)prog_len[pg] * sites, {(int)i, pg, blk, first[v++] + q}});
            }
        }
    }
    auto by_cost = [](const Cand& x, const Cand& y) { return x.cost > y.cost; };
    std::stable_sort(c1.begin(), c1.end(), by_cost);
    std::stable_sort(c2.begin(), c2.end(), by_cost);
    b->items1.clear();
    b->items2.clear();
    for (auto& x : c1) b->items1.push_back(x.w);
    // Workgroup g runs on XCD g mod 8 (workgroups are dealt to the XCDs in
    // turn) and the grid is a multiple of 8, so item j runs on XCD j mod 8.
    // The variants of one (alignment, site block) are dealt to one XCD, in
    // cost order there, so those running side by side share the block's
    // stored partials in that XCD's L2.
    std::vector<std::vector<WorkItem>> q(8);
    std::unordered_map<long, int> xcd_of;
    for (const auto& x : c2) {
        const long key = (long)x.w.inst << 32 | (unsigned)x.w.block;
        const auto it = xcd_of.emplace(key, (int)(xcd_of.size() % 8)).first;
        q[it->second].push_back(x.w);
    }
    size_t next[8] = {};
    for (size_t j = 0; j < c2.size(); ++j) {
        int x = (int)(j % 8);
        if (next[x] == q[x].size()) {    // this XCD's share is done: take from the longest remaining
            for (int k = 0; k < 8; ++k)
                if (q[k].size() - next[k] > q[x].size() - next[x]) x = k;
        }
        b->items2.push_back(q[x][next[x]++]);
    }
    for (auto& s : b->seg2) s += base2;
    b->seg2.front() = base2;
}

int upload_work_lists(phylo_batch* b) {
    b->n_items1 = (int)b->items1.size();
    b->n_items2 = (int)b->items2.size();
    b->n_seg1 = (int)b->seg1.size() - 1;
    b->n_seg2 = (int)b->seg2.size() - 1;
    auto up = [&](auto*& dptr, const auto& vec) -> hipError_t {
        if (dptr) { hipError_t e = hipFree(dptr); if (e != hipSuccess) return e; dptr = nullptr; }
        hipError_t e = hipMalloc(reinterpret_cast<void**>(&dptr), std::max<size_t>(vec.size(), 1) * sizeof(vec[0]));
        if (e != hipSuccess) return e;
        if (vec.empty()) return hipSuccess;
        return hipMemcpy(dptr, vec.data(), vec.size() * sizeof(vec[0]), hipMemcpyHostToDevice);
    };
    HIPCHK(b, up(b->d_items1, b->items1));
    HIPCHK(b, up(b->d_items2, b->items2));
    HIPCHK(b, up(b->d_seg1, b->seg1));
    HIPCHK(b, up(b->d_seg2, b->seg2));
    if (b->d_partial) HIPCHK(b, hipFree(b->d_partial));
    b->d_partial = nullptr;
    HIPCHK(b, hipMalloc(&b->d_partial, std::max<size_t>(b->n_items1 + b->n_items2, 1) * sizeof(double)));
    return PHYLO_OK;
}

// Enqueues one full batch run (both phases and the segment sums) on the stream.
int enqueue_batch_run(phylo_batch* b) {
    int rc = enqueue_batch_phase(b, 1, b->blocks_per_item);
    if (rc) return rc;
    rc = enqueue_batch_phase(b, 2, b->blocks_per_item);
    if (rc) return rc;
    if (b->n_seg1)
        hipLaunchKernelGGL(segment_sum_kernel, dim3((b->n_seg1 + 63) / 64), dim3(64), 0, b->stream, b->d_partial,
                           b->d_seg1, b->n_seg1, b->d_out);
    if (b->n_seg2)
        hipLaunchKernelGGL(segment_sum_kernel, dim3((b->n_seg2 + 63) / 64), dim3(64), 0, b->stream, b->d_partial,
                           b->d_seg2, b->n_seg2, b->d_out + b->n_seg1);
    HIPCHK(b, hipGetLastError());
    HIPCHK(b, hipMemcpyAsync(b->h_out, b->d_out, (size_t)(b->n_seg1 + b->n_seg2) * sizeof(double),
                             hipMemcpyDeviceToHost, b->stream));
    return PHYLO_OK;
}

}  // namespace

extern "C" {

int phylo_batch_create(int device, int n_cat, int precision, phylo_batch** out) {
    if (!out) return PHYLO_ERR_ARG;
    *out = nullptr;
    if (n_cat != 1 && n_cat != 2 && n_cat != 4 && n_cat != 8) return PHYLO_ERR_ARG;
    if (precision != PHYLO_PREC_FP32 && precision != PHYLO_PREC_FP64) return PHYLO_ERR_ARG;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return PHYLO_ERR_HIP;
    if (device < 0 || device >= n) return PHYLO_ERR_ARG;
    phylo_batch* b = new phylo_batch();
    b->device = device;
    b->C = n_cat;
    b->cat_waves = n_cat;
    b->prec = precision;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess) {
        delete b;
        return PHYLO_ERR_HIP;
    }
    *out = b;
    return PHYLO_OK;
}

const char* phylo_batch_last_error(const phylo_batch* b) { return b ? b->err.c_str() : "null batch"; }

int phylo_batch_add(phylo_batch* b, int n_taxa, int n_sites, int n_states, const uint32_t* masks, int n_nodes, int root,
                    const int32_t* left, const int32_t* right, const double* eff, const double* Q,
                    const double* root_w, const double* pi, const double* cat_rates, const double* cat_w,
                    int nni_mode, int* index) {
    if (!b) return PHYLO_ERR_ARG;
    if (b->finalized) return fail(b, PHYLO_ERR_STATE, "batch_add after batch_finalize");
    if (n_states < 1 || n_states > 4) return fail(b, PHYLO_ERR_UNSUPPORTED, "batch_add: batches hold up to 4 states");
    if (n_taxa < 2 || n_sites < 1 || n_nodes <= n_taxa || root < n_taxa || root >= n_nodes || !masks || !left ||
        !right || !eff || !Q || !root_w || !pi || !cat_rates || !cat_w || nni_mode < -1 || nni_mode > 1)
        return fail(b, PHYLO_ERR_ARG, "batch_add: bad arguments");
    phylo_batch::Inst I;
    I.n_taxa = n_taxa;
    I.m = n_sites;
    I.m_pad = (n_sites + 63) / 64 * 64;
    I.S = n_states;
    I.n_nodes = n_nodes;
    I.root = root;
    I.left.assign(left, left + n_nodes);
    I.right.assign(right, right + n_nodes);
    std::vector<int> indeg(n_nodes, 0);
    for (int v = 0; v < n_nodes; ++v) {
        if (v < n_taxa && (left[v] != -1 || right[v] != -1)) return fail(b, PHYLO_ERR_ARG, "batch_add: tip with children");
        if (v >= n_taxa) {
            if (left[v] < 0 || left[v] >= n_nodes || right[v] < -1 || right[v] >= n_nodes)
                return fail(b, PHYLO_ERR_ARG, "batch_add: bad child index");
            ++indeg[left[v]];
            if (right[v] >= 0) ++indeg[right[v]];
        }
    }
    for (int v = 0; v < n_nodes; ++v)
        if ((v == root) != (indeg[v] == 0) || indeg[v] > 1) return fail(b, PHYLO_ERR_ARG, "batch_add: not a tree");
    const uint32_t valid = (1u << n_states) - 1u;
    I.masks.assign(masks, masks + (size_t)n_taxa * n_sites);
    for (auto& x : I.masks) {
        x &= valid;
        if (!x) return fail(b, PHYLO_ERR_DATA, "batch_add: a tip lacks a character state");
    }
    // P per (node, category) from the reversible eigen-system, on the host
    std::vector<double> V, lam, Vinv;
    if (!phylo_math::is_reversible(n_states, Q, pi) || !phylo_math::reversible_eigen(n_states, Q, pi, V, lam, Vinv))
        return fail(b, PHYLO_ERR_ARG, "batch_add: Q must be reversible with positive pi");
    I.P.assign((size_t)n_nodes * b->C * 16, 0.0);
    for (int v = 0; v < n_nodes; ++v) {
        if (v == root) continue;
        if (!std::isfinite(eff[v])) return fail(b, PHYLO_ERR_ARG, "batch_add: non-finite branch length");
        for (int k = 0; k < b->C; ++k) {
            const double t = eff[v] * cat_rates[k];
            double* dst = I.P.data() + ((size_t)v * b->C + k) * 16;
            for (int i = 0; i < n_states; ++i)
                for (int j = 0; j < n_states; ++j) {
                    double a = 0.0;
                    for (int q = 0; q < n_states; ++q)
                        a += V[(size_t)i * n_states + q] * std::exp(lam[q] * t) * Vinv[(size_t)q * n_states + j];
                    dst[i * 4 + j] = a;
                }
        }
    }
    for (double& r : I.roots) r = 0.0;
    for (int i = 0; i < n_states; ++i) I.roots[i] = root_w[i];
    for (int k = 0; k < b->C; ++k) I.roots[kRootCatOff + k] = cat_w[k];
    I.nni_mode = nni_mode;
    b->inst.push_back(std::move(I));
    if (index) *index = (int)b->inst.size() - 1;
    return PHYLO_OK;
}

int phylo_batch_finalize(phylo_batch* b, int autotune) {
    if (!b) return PHYLO_ERR_ARG;
    if (b->finalized) return fail(b, PHYLO_ERR_STATE, "batch already finalized");
    if (b->inst.empty()) return fail(b, PHYLO_ERR_STATE, "empty batch");
    HIPCHK(b, hipSetDevice(b->device));
    std::vector<Op> ops;
    std::vector<int> prog_begin{0}, prog_len;
    long tips_bytes = 0, clv_elems = 0, scl_elems = 0, P_elems = 0;
    int roots_elems = 0;
    b->updates1 = b->updates2 = 0;
    b->var_offsets.assign(1, 0);
    b->var_prog_of.clear();
    for (auto& I : b->inst) {
        I.tips_off = tips_bytes;
        tips_bytes += (long)I.n_taxa * I.m_pad;
        I.clv_off = clv_elems;
        clv_elems += (long)(I.n_nodes - I.n_taxa) * b->C * I.m_pad * 4;
        I.scl_off = scl_elems;
        scl_elems += (long)(I.n_nodes - I.n_taxa) * b->C * I.m_pad;
        I.P_off = P_elems;
        P_elems += (long)I.n_nodes * b->C * 16;
        I.roots_off = roots_elems;
        roots_elems += kMaxStates + kMaxCat;
        const size_t before = ops.size();
        if (!build_full_ops(&I, ops)) return fail(b, PHYLO_ERR_ARG, "batch: internal nodes unreachable from the root");
        if (I.nni_mode < 0) drop_unread_stores(ops, before, ops.size());   // no variant reads the partials
        I.full_prog = (int)prog_begin.size() - 1;
        prog_begin.push_back((int)ops.size());
        prog_len.push_back((int)(ops.size() - before));
        I.full_ops = (long)(ops.size() - before);
        b->updates1 += (double)I.full_ops * I.m * b->C;
        I.nni_prog.clear();
        I.nni_node.clear();
        I.nni_side.clear();
        I.nni_ops = 0;
        if (I.nni_mode >= 0) {
            const std::vector<int> par = parents_of(&I);
            enumerate_nni(&I, I.nni_mode, I.nni_node, I.nni_side);
            for (size_t k = 0; k < I.nni_node.size(); ++k) {
                const size_t b0 = ops.size();
                if (append_nni_program(&I, par, I.nni_node[k], I.nni_side[k], ops)) {
                    I.nni_prog.push_back((int)prog_begin.size() - 1);
                    prog_begin.push_back((int)ops.size());
                    prog_len.push_back((int)(ops.size() - b0));
                    I.nni_ops += (long)(ops.size() - b0);
                } else {
                    I.nni_prog.push_back(-1);
                }
            }
            b->updates2 += (double)I.nni_ops * I.m * b->C;
        }
        b->var_offsets.push_back(b->var_offsets.back() + (int)I.nni_node.size());
    }
    b->n_var = b->var_offsets.back();
    if (!root_ops_last(ops, prog_begin)) return fail(b, PHYLO_ERR_STATE, "internal: root op not last");
    // device buffers
    const size_t elt = b->prec == PHYLO_PREC_FP64 ? sizeof(double) : sizeof(float);
    HIPCHK(b, hipMalloc(&b->d_tips, std::max<long>(tips_bytes, 1)));
    HIPCHK(b, hipMalloc(&b->d_clv, std::max<size_t>((size_t)clv_elems * elt, 16)));
    HIPCHK(b, hipMalloc(&b->d_scl, std::max<size_t>((size_t)scl_elems * sizeof(int), 16)));
    HIPCHK(b, hipMalloc(&b->d_P, std::max<size_t>((size_t)P_elems * sizeof(double), 16)));
    HIPCHK(b, hipMalloc(&b->d_roots, (size_t)roots_elems * sizeof(double)));
    HIPCHK(b, hipMalloc(&b->d_ops, std::max<size_t>(ops.size(), 1) * sizeof(Op)));
    HIPCHK(b, hipMalloc(&b->d_prog, prog_begin.size() * sizeof(int)));
    HIPCHK(b, hipMalloc(&b->d_inst, b->inst.size() * sizeof(BatchInst)));
    std::vector<BatchInst> hinst;
    std::vector<uint8_t> codes;
    for (auto& I : b->inst) {
        BatchInst d{};
        d.tips_off = I.tips_off;
        d.clv_off = I.clv_off;
        d.scl_off = I.scl_off;
        d.P_off = I.P_off;
        d.roots_off = I.roots_off;
        d.m = I.m;
        d.m_pad = I.m_pad;
        hinst.push_back(d);
        codes.assign((size_t)I.n_taxa * I.m_pad, 0x0F);
        for (int t = 0; t < I.n_taxa; ++t)
            for (int s = 0; s < I.m; ++s) codes[(size_t)t * I.m_pad + s] = (uint8_t)I.masks[(size_t)t * I.m + s];
        HIPCHK(b, hipMemcpy(b->d_tips + I.tips_off, codes.data(), codes.size(), hipMemcpyHostToDevice));
        HIPCHK(b, hipMemcpy(b->d_P + I.P_off, I.P.data(), I.P.size() * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(b, hipMemcpy(b->d_roots + I.roots_off, I.roots, sizeof(I.roots), hipMemcpyHostToDevice));
        std::vector<uint32_t>().swap(I.masks);
        std::vector<double>().swap(I.P);
    }
    HIPCHK(b, hipMemcpy(b->d_inst, hinst.data(), hinst.size() * sizeof(BatchInst), hipMemcpyHostToDevice));
    order_children(ops);
    HIPCHK(b, hipMemcpy(b->d_ops, ops.data(), ops.size() * sizeof(Op), hipMemcpyHostToDevice));
    HIPCHK(b, hipMemcpy(b->d_prog, prog_begin.data(), prog_begin.size() * sizeof(int), hipMemcpyHostToDevice));
    const int n_out = (int)b->inst.size() + b->n_var;
    HIPCHK(b, hipMalloc(&b->d_out, std::max(n_out, 1) * sizeof(double)));
    HIPCHK(b, hipHostMalloc(reinterpret_cast<void**>(&b->h_out), std::max(n_out, 1) * sizeof(double)));
    // autotune the work-item size (64-site blocks per item) with the
    // categories on separate waves, then the one-wave category tiling around
    // the best size
    struct Cfg { int cw, bpi, cap; };
    std::vector<Cfg> cands;
    if (autotune) {
        for (int bpi : {1, 2, 4, 8}) cands.push_back({b->C, bpi, 0});
    } else {
        cands.push_back({b->cat_waves, b->blocks_per_item, b->per_cu_cap});
    }
    float best = 1e30f;
    Cfg best_cfg = cands.front();
    auto trial = [&](const Cfg& cfg) -> int {
        b->cat_waves = cfg.cw;
        b->per_cu_cap = cfg.cap;
        int rc = batch_grid(b);
        if (rc) return rc;
        build_work_lists(b, cfg.bpi, prog_len);
        rc = upload_work_lists(b);
        if (rc) return rc;
        b->blocks_per_item = cfg.bpi;
        float ms = 0.f;
        for (int rep = 0; rep < (autotune ? 2 : 0); ++rep) {
            HIPCHK(b, hipEventRecord(b->ev0, b->stream));
            rc = enqueue_batch_run(b);
            if (rc) return rc;
            HIPCHK(b, hipEventRecord(b->ev1, b->stream));
            HIPCHK(b, hipStreamSynchronize(b->stream));
            HIPCHK(b, hipEventElapsedTime(&ms, b->ev0, b->ev1));
        }
        if (autotune && ms < best) { best = ms; best_cfg = cfg; }
        return PHYLO_OK;
    };
    for (const Cfg& cfg : cands) {
        int rc = trial(cfg);
        if (rc) return rc;
    }
    if (autotune && b->C > 1) {
        const int bpi0 = best_cfg.bpi;
        for (int bpi : {bpi0, 2 * bpi0}) {
            int rc = trial({1, bpi, 0});
            if (rc) return rc;
        }
    }
    if (autotune) {
        // fewer resident workgroups per CU: the NNI programs mostly read
        // stored partials, and more waves in flight can thrash the caches
        b->cat_waves = best_cfg.cw;
        int rc = batch_grid(b);
        if (rc) return rc;
        const Cfg base = best_cfg;
        for (int cap = b->max_per_cu - 1; cap >= std::max(1, b->max_per_cu - 3); --cap) {
            rc = trial({base.cw, base.bpi, cap});
            if (rc) return rc;
        }
    }
    if (b->cat_waves != best_cfg.cw || b->blocks_per_item != best_cfg.bpi || b->per_cu_cap != best_cfg.cap) {
        b->cat_waves = best_cfg.cw;
        b->per_cu_cap = best_cfg.cap;
        int rc = batch_grid(b);
        if (rc) return rc;
        build_work_lists(b, best_cfg.bpi, prog_len);
        rc = upload_work_lists(b);
        if (rc) return rc;
        b->blocks_per_item = best_cfg.bpi;
    }
    // capture the whole run as a graph, replayed by phylo_batch_run
    HIPCHK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_batch_run(b);
    hipGraph_t g = nullptr;
    hipError_t ce = hipStreamEndCapture(b->stream, &g);
    if (rc) return rc;
    HIPCHK(b, ce);
    b->graph = g;
    HIPCHK(b, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    b->finalized = true;
    return PHYLO_OK;
}

int phylo_batch_run(phylo_batch* b, double* totals, double* nni_lnl) {
    if (!b) return PHYLO_ERR_ARG;
    if (!b->finalized) return fail(b, PHYLO_ERR_STATE, "batch_run before batch_finalize");
    HIPCHK(b, hipSetDevice(b->device));
    HIPCHK(b, hipEventRecord(b->ev0, b->stream));
    HIPCHK(b, hipGraphLaunch(b->exec, b->stream));
    HIPCHK(b, hipEventRecord(b->ev1, b->stream));
    HIPCHK(b, hipStreamSynchronize(b->stream));
    float ms = 0.f;
    HIPCHK(b, hipEventElapsedTime(&ms, b->ev0, b->ev1));
    b->last_ms = ms;
    const int ni = (int)b->inst.size();
    if (totals)
        for (int i = 0; i < ni; ++i) totals[i] = b->h_out[i];
    if (nni_lnl) {
        int seg = 0;
        for (int i = 0; i < ni; ++i) {
            const auto& I = b->inst[i];
            for (size_t k = 0; k < I.nni_prog.size(); ++k) {
                double v = b->h_out[i];                  // no-op variants keep the base tree
                if (I.nni_prog[k] >= 0) v = b->h_out[ni + seg++];
                nni_lnl[b->var_offsets[i] + k] = v;
            }
        }
    }
    return PHYLO_OK;
}

int phylo_batch_nni_layout(phylo_batch* b, int32_t* offsets, int32_t* node, int32_t* side) {
    if (!b) return PHYLO_ERR_ARG;
    if (!b->finalized) return fail(b, PHYLO_ERR_STATE, "batch_nni_layout before batch_finalize");
    for (size_t i = 0; i < b->var_offsets.size() && offsets; ++i) offsets[i] = b->var_offsets[i];
    for (size_t i = 0; i < b->inst.size(); ++i)
        for (size_t k = 0; k < b->inst[i].nni_node.size(); ++k) {
            if (node) node[b->var_offsets[i] + k] = b->inst[i].nni_node[k];
            if (side) side[b->var_offsets[i] + k] = b->inst[i].nni_side[k];
        }
    return PHYLO_OK;
}

int phylo_batch_stats(phylo_batch* b, double* updates_full, double* updates_nni, double* last_ms,
                      int* sites_per_item, int* n_variants) {
    if (!b) return PHYLO_ERR_ARG;
    if (updates_full) *updates_full = b->updates1;
    if (updates_nni) *updates_nni = b->updates2;
    if (last_ms) *last_ms = b->last_ms;
    if (sites_per_item) *sites_per_item = 64 * b->blocks_per_item;
    if (n_variants) *n_variants = b->n_var;
    return PHYLO_OK;
}

int phylo_batch_set_tiling(phylo_batch* b, int sites_per_item, int category_waves) {
    if (!b) return PHYLO_ERR_ARG;
    if (b->finalized) return fail(b, PHYLO_ERR_STATE, "batch already finalized");
    if (sites_per_item < 64 || sites_per_item > 4096 || sites_per_item % 64)
        return fail(b, PHYLO_ERR_ARG, "set_tiling: sites per item must be a multiple of 64 up to 4096");
    if (category_waves != 1 && category_waves != b->C)
        return fail(b, PHYLO_ERR_ARG, "set_tiling: category waves must be 1 or the number of categories");
    b->blocks_per_item = sites_per_item / 64;
    b->cat_waves = category_waves;
    return PHYLO_OK;
}

int phylo_batch_tiling(phylo_batch* b, int* sites_per_item, int* category_waves, int* workgroups_per_cu) {
    if (!b) return PHYLO_ERR_ARG;
    if (sites_per_item) *sites_per_item = 64 * b->blocks_per_item;
    if (category_waves) *category_waves = b->cat_waves;
    if (workgroups_per_cu)
        *workgroups_per_cu = b->per_cu_cap > 0 ? std::min(b->per_cu_cap, b->max_per_cu) : b->max_per_cu;
    return PHYLO_OK;
}

void phylo_batch_destroy(phylo_batch* b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    void* bufs[] = {b->d_tips, b->d_clv, b->d_scl, b->d_P, b->d_roots, b->d_ops, b->d_prog, b->d_inst,
                    b->d_items1, b->d_items2, b->d_partial, b->d_seg1, b->d_seg2, b->d_out};
    for (void* p : bufs)
        if (p) (void)hipFree(p);
    if (b->h_out) (void)hipHostFree(b->h_out);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

}  // extern "C"
