// kernels.h -- device kernels of libmeshprop (prologue, propagation, dp reduction) and the
// launch plan shared by the host translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <atomic>
#include <type_traits>

#include "crossing.h"
#include "meshprop.h"
#include "propagate.h"

namespace mp {

// =======================================================================================
// Prologue: row tables
// =======================================================================================
enum {
    TAB_FWD = 1,
    TAB_DIFF = 2,
    TAB_INV = 4,
    TAB_ADJ = 8,
    TAB_DTDP = 16,
    TAB_DIFF_ADJ = 32,
    TAB_INV_B = 64,
    TAB_KUNI = 128,
    TAB_SCALE = 256  // per-waveguide |det T| of scaled-unitary crossings (1 on identity rows)
};

// Slot and role of waveguide w in layer l.  Returns whether an active crossing acts on it.
// NN: gpu/gmem.cu:236-240 (slot m couples 2m + shift%2, +1; active for
// shift/2 <= m < shift/2 + len).  FFT: gpu/test.py:73-77.
__device__ inline bool wg_role(const int fft, const int* g0, const int* g1, const int l, const int w, const int N,
                               int& m, bool& upper) {
    if (w >= N) return false;
    if (fft) {
        const int st = g0[l];
        if (st < 0 || st > 20) return false;
        const int S = 1 << st;
        upper = (w & S) == 0;
        const int i = upper ? w : w - S;
        if (i + S >= N) return false;
        m = (i / (2 * S)) * S + (i % (2 * S));
        return true;
    }
    const int sh = g1[l], ln = g0[l];
    const int par = sh & 1;
    if (w < par) return false;
    const int q = w - par;
    m = q >> 1;
    upper = (q & 1) == 0;
    const int partner = upper ? w + 1 : w - 1;
    return (m >= (sh >> 1)) && (m < (sh >> 1) + ln) && (partner < N) && (partner >= 0);
}

__device__ inline int layer_kind(const int fft, const int* g0, const int* g1, const int l, const int NP) {
    if (fft) {
        const int st = g0[l];
        if (st < 0 || st > 20 || (2 << st) > NP) return KIND_IDENTITY;
        return 2 + st;
    }
    return g1[l] & 1;
}

template <class R>
__device__ __forceinline__ R make_row(const dc a, const dc b);
template <>
__device__ __forceinline__ float4 make_row<float4>(const dc a, const dc b) {
    return make_float4((float)a.r, (float)a.i, (float)b.r, (float)b.i);
}
template <>
__device__ __forceinline__ float2 make_row<float2>(const dc a, const dc b) {
    return make_float2((float)a.r, (float)b.r);
}

// Row (coefficient on self, coefficient on partner) of matrix M for the upper / lower waveguide.
__device__ __forceinline__ void row_of(const M2& M, const bool upper, dc& a, dc& b) {
    if (upper) {
        a = M.a00;
        b = M.a01;
    } else {
        a = M.a11;
        b = M.a10;
    }
}

// Row tables are stored k-major inside each layer: row of waveguide w = t*W + k (lane t of a
// lane group, k-th waveguide of that lane) sits at [l][k][t], so the k-th row load of all
// lanes reads one contiguous block.

// dT/dp_k of every slot (zero for inactive slots), for the in-kernel dJdT -> dp step.
// Layout: dtdp[(l*np*4 + k*4 + e)*NPh + m] with e = 00, 01, 10, 11 (slot-contiguous), element
// type float2 (complex) or float (real).
template <class T>
__device__ __forceinline__ T dc_to(const dc z);
template <>
__device__ __forceinline__ float2 dc_to<float2>(const dc z) {
    return make_float2((float)z.r, (float)z.i);
}
template <>
__device__ __forceinline__ float dc_to<float>(const dc z) {
    return (float)z.r;
}

// K-form coefficients of unitary crossings (TAB_KUNI, written into the dT/dp buffer):
// K_k = T^-1 dT/dp_k (= T^H dT/dp_k for unitary T; the same for a scaled unitary T = c U) is
// anti-Hermitian, [[i a, c], [-conj(c), i d]], stored per (layer,
// parameter, slot) as float4 (a, d, c.re, c.im); for real (orthogonal) crossings K_k =
// [[0, c], [-c, 0]] and the float c is stored.  Layout: ktab[(l*np + k)*NPh + m].
template <class T>
struct KOf;
template <>
struct KOf<float2> {
    typedef float4 type;
};
template <>
struct KOf<float> {
    typedef float type;
};
// Conversion coefficients per slot: the K-form (one element per parameter) or dT/dp (four
// entries per parameter); both tables index [(l*NC + e)*NPh + m].
template <class T, bool KF>
struct CoefOf {
    typedef T type;
    static constexpr int per_param = 4;
};
template <class T>
struct CoefOf<T, true> {
    typedef typename KOf<T>::type type;
    static constexpr int per_param = 1;
};

// Field element type of a row type (float4 rows: complex64 fields; float2 rows: real fields).
template <class R>
struct FieldOfRow;
template <>
struct FieldOfRow<float4> {
    typedef float2 type;
};
template <>
struct FieldOfRow<float2> {
    typedef float type;
};

// Stage order of the NN-specialised walks: lane t of a 64-lane wave owns kso consecutive slots
// t*kso .. t*kso + kso - 1, and its slot p sits at position p*64 + t of the wave's block of
// 64*kso positions; a vector split over several waves (lane index over the whole vector
// >= 64) takes one block per wave, in order.  kso == 0: slot order.
__host__ __device__ inline int stage_pos(const int m, const int kso) {
    if (kso <= 0) return m;
    const int ln = m / kso;
    return (ln / 64) * (64 * kso) + (m % kso) * 64 + ln % 64;
}
__host__ __device__ inline int stage_slot(const int pos, const int kso) {
    if (kso <= 0) return pos;
    const int r = pos % (64 * kso);
    return (pos - r) + (r % 64) * kso + r / 64;
}

// One thread per (layer, slot m < NP/2): the crossing of the slot is evaluated once (FP64) and
// both of its waveguides' rows are written to every requested table, with its dT/dp for
// TAB_DTDP; waveguides that no active crossing touches get identity rows (and inactive slots
// a zero dT/dp).  Every waveguide of the padded width is written by exactly one thread: slot m
// holds waveguides (i, j) = (2m + shift%2, +1) in NN layers (slot 0 of an odd-shift layer also
// writes waveguide 0) and (i, i + 2^s) in FFT layers (all waveguides identity when the
// stride does not fit the padded width).
template <class R>
__device__ void slot_tables_at(const int X, const int N, const int NP, const int LPV, const int fft,
                               const int* __restrict__ g0, const int* __restrict__ g1, const float* __restrict__ p,
                               const int ldp, const float* __restrict__ s, const int lds,
                               const float* __restrict__ dp, const int tabs, int* __restrict__ kind,
                               R* __restrict__ tab_a, R* __restrict__ tab_b,
                               typename FieldOfRow<R>::type* __restrict__ dtdp, float* __restrict__ scl,
                               const int l, const int m, const int kso = 0) {
    typedef typename FieldOfRow<R>::type F;
    const int kd = layer_kind(fft, g0, g1, l, NP);
    if (m == 0 && kind) kind[l] = kd;
    const int NPh = NP / 2;
    // Index of slot m in the K-form table: slot order, or (kso > 0) the NN walks' stage order.
    const int km = stage_pos(m, kso);
    int i, j;
    bool act;
    if (fft) {
        if (kd == KIND_IDENTITY) {
            i = 2 * m;
            j = i + 1;
            act = false;
        } else {
            const int S = 1 << (kd - 2);
            i = 2 * S * (m / S) + (m % S);
            j = i + S;
            act = j < N;
        }
    } else {
        const int sh = g1[l], ln = g0[l];
        i = 2 * m + (sh & 1);
        j = i + 1;
        act = (m >= (sh >> 1)) && (m < (sh >> 1) + ln) && (j < N);
    }
    const int Wd = NP / LPV;
    auto pos = [&](const int w) { return (size_t)l * NP + (size_t)(w % Wd) * LPV + w / Wd; };
    const dc one = mk(1, 0), zero = mk(0, 0);
    auto put = [&](R* tab, const int w, const dc a, const dc b) {
        if (w < NP) tab[pos(w)] = make_row<R>(a, b);
    };
    auto put_scale = [&](const int w, const float v) {
        if (w < NP) scl[pos(w)] = v;
    };
    const int np = n_phase_of(X), ns = n_split_of(X);
    if (!fft && (g1[l] & 1) && m == 0) {  // waveguide 0 of an odd-shift layer: no crossing
        if (tabs & (TAB_FWD | TAB_INV)) put(tab_a, 0, one, zero);
        if (tabs & (TAB_DIFF | TAB_DIFF_ADJ)) put(tab_b, 0, zero, zero);
        if (tabs & (TAB_ADJ | TAB_INV_B)) put(tab_b, 0, one, zero);
        if (tabs & TAB_SCALE) put_scale(0, 1.f);
    }
    if (!act) {
        for (int w = i; w <= j; w += j - i) {
            if (tabs & (TAB_FWD | TAB_INV)) put(tab_a, w, one, zero);
            if (tabs & (TAB_DIFF | TAB_DIFF_ADJ)) put(tab_b, w, zero, zero);
            if (tabs & (TAB_ADJ | TAB_INV_B)) put(tab_b, w, one, zero);
            if (tabs & TAB_SCALE) put_scale(w, 1.f);
        }
        if (tabs & TAB_DTDP) {
            F* out = dtdp + (size_t)l * NPh * np * 4 + m;
            for (int e = 0; e < np * 4; ++e) out[(size_t)e * NPh] = dc_to<F>(zero);
        }
        if (tabs & TAB_KUNI) {
            typedef typename KOf<F>::type K;
            K* out = reinterpret_cast<K*>(dtdp) + (size_t)l * np * NPh + km;
            for (int k2 = 0; k2 < np; ++k2) out[(size_t)k2 * NPh] = K{};
        }
        return;
    }
    const float* pp = p + (size_t)l * ldp + (size_t)m * np;
    const float* sp = (s && ns) ? s + (size_t)l * lds + (size_t)m * ns : nullptr;
    M2 T, dT[2];
    const bool need_d = (tabs & (TAB_DIFF | TAB_DIFF_ADJ | TAB_DTDP | TAB_KUNI)) != 0;
    crossing_eval(X, pp, sp, &T, need_d ? dT : nullptr);
    // Rows of matrix M: upper waveguide i (M00, M01), lower waveguide j (M11, M10).
    auto put2 = [&](R* tab, const M2& M) {
        put(tab, i, M.a00, M.a01);
        put(tab, j, M.a11, M.a10);
    };
    if (tabs & TAB_FWD) put2(tab_a, T);
    if (tabs & (TAB_DIFF | TAB_DIFF_ADJ)) {
        const float* dpp = dp + (size_t)l * ldp + (size_t)m * np;
        M2 D = mscale(dT[0], (double)dpp[0]);
        if (np > 1) D = madd(D, mscale(dT[1], (double)dpp[1]));
        if (tabs & TAB_DIFF) put2(tab_b, D);
        if (tabs & TAB_DIFF_ADJ) put2(tab_b, mherm(D));
    }
    if (tabs & (TAB_INV | TAB_INV_B)) {
        // Unitary crossings: T^-1 = T^H.  Otherwise the adjugate inverse.
        const M2 Ti = (X == X_GENERIC) ? minv(T) : mherm(T);
        if (tabs & TAB_INV) put2(tab_a, Ti);
        if (tabs & TAB_INV_B) put2(tab_b, Ti);
    }
    if (tabs & TAB_ADJ) put2(tab_b, mherm(T));
    if (tabs & TAB_DTDP) {
        F* out = dtdp + (size_t)l * NPh * np * 4 + m;
        for (int k2 = 0; k2 < np; ++k2) {
            out[(size_t)(k2 * 4 + 0) * NPh] = dc_to<F>(dT[k2].a00);
            out[(size_t)(k2 * 4 + 1) * NPh] = dc_to<F>(dT[k2].a01);
            out[(size_t)(k2 * 4 + 2) * NPh] = dc_to<F>(dT[k2].a10);
            out[(size_t)(k2 * 4 + 3) * NPh] = dc_to<F>(dT[k2].a11);
        }
    }
    if (tabs & TAB_SCALE) {
        // Scaled unitary T = c U: T^H T = |det T| I, and the adjoint rows are |det T| times
        // the inverse rows.  A crossing that is not (beyond FP64 rounding) gets NaN.
        const M2 E = mmul(mherm(T), T);
        const double sc = cabs(mdet(T));
        const double err = fabs(E.a00.r - sc) + fabs(E.a11.r - sc) + fabs(E.a00.i) + fabs(E.a11.i) +
                           cabs(E.a01) + cabs(E.a10);
        const float v = (err <= 1e-9 * sc) ? (float)sc : __int_as_float(0x7fc00000);
        put_scale(i, v);
        put_scale(j, v);
    }
    if (tabs & TAB_KUNI) {
        // Anti-Hermitian part of K = T^-1 dT/dp (T^-1 = T^H for a unitary T; for a scaled
        // unitary T = c U it equals U^H dU/dp).
        typedef typename KOf<F>::type K;
        K* out = reinterpret_cast<K*>(dtdp) + (size_t)l * np * NPh + km;
        const M2 Ti = (X == X_GENERIC) ? minv(T) : mherm(T);
        for (int k2 = 0; k2 < np; ++k2) {
            const M2 Kp = mmul(Ti, dT[k2]);
            const double cr = 0.5 * (Kp.a01.r - Kp.a10.r), ci = 0.5 * (Kp.a01.i + Kp.a10.i);
            if constexpr (sizeof(K) == sizeof(float4))
                out[(size_t)k2 * NPh] = K{(float)Kp.a00.i, (float)Kp.a11.i, (float)cr, (float)ci};
            else
                out[(size_t)k2 * NPh] = K((float)cr);
        }
    }
}

// Prologue of one call: row tables (tabs) and, with TAB_DTDP, the dT/dp table of every slot,
// all in one launch, one thread per (layer, slot).  kso > 0: K-form table in stage order.
template <class R>
__global__ void k_rows(int X, int N, int L, int NP, int LPV, int fft, const int* __restrict__ g0,
                       const int* __restrict__ g1,
                       const float* __restrict__ p, int ldp, const float* __restrict__ s, int lds,
                       const float* __restrict__ dp, int tabs, int* __restrict__ kind, R* __restrict__ tab_a,
                       R* __restrict__ tab_b, typename FieldOfRow<R>::type* __restrict__ dtdp,
                       float* __restrict__ scl, int kso) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    const int l = blockIdx.y;
    if (l >= L || m >= NP / 2) return;
    slot_tables_at<R>(X, N, NP, LPV, fft, g0, g1, p, ldp, s, lds, dp, tabs, kind, tab_a, tab_b, dtdp, scl, l, m,
                      kso);
}

// =======================================================================================
// Propagation kernels
// =======================================================================================
// Layer kinds seen through the constant address space: read-only for the whole launch, so
// their loads stay scalar (s_load through the constant cache) also where the compiler cannot
// prove that no store of the kernel reaches them (after a call, or beside global stores).
typedef __attribute__((address_space(4))) const int ConstInt;

struct Lanes {
    int lane, wave, grp, t;
    long vbase;
};
template <int V>
__device__ __forceinline__ Lanes lanes_of(const int LPV, const long tile) {
    Lanes q;
    q.lane = threadIdx.x & 63;
    q.wave = threadIdx.x >> 6;
    const int G = 64 / LPV;
    q.grp = q.lane / LPV;
    q.t = q.lane - q.grp * LPV;
    const int nw = blockDim.x >> 6;
    q.vbase = ((tile * nw + q.wave) * G + q.grp) * V;
    return q;
}

template <class T, int W, int V>
__device__ __forceinline__ void load_tile(T (&u)[V][W], const T* src, const Lanes& q, const int N, const int B,
                                          const int ldu) {
#pragma unroll
    for (int v = 0; v < V; ++v) {
        const long b = q.vbase + v;
#pragma unroll
        for (int k = 0; k < W; ++k) {
            const int w = q.t * W + k;
            u[v][k] = (b < B && w < N) ? src[b * ldu + w] : fzero(T());
        }
    }
}

template <class T, int W, int V>
__device__ __forceinline__ void store_tile(const T (&u)[V][W], T* dst, const Lanes& q, const int N, const int B,
                                           const int ldu) {
#pragma unroll
    for (int v = 0; v < V; ++v) {
        const long b = q.vbase + v;
#pragma unroll
        for (int k = 0; k < W; ++k) {
            const int w = q.t * W + k;
            if (b < B && w < N) dst[b * ldu + w] = u[v][k];
        }
    }
}

template <class R, int W>
__device__ __forceinline__ void load_rows(R (&r)[W], const R* base, const int LPV) {
#pragma unroll
    for (int k = 0; k < W; ++k) r[k] = base[(size_t)k * LPV];
}

// A lane's entry of a k-major table: a wave-uniform base (layer offsets are added to it in
// scalar registers) and the lane's byte offset, a 32-bit value fixed for the whole walk.  Rows
// are read with raw buffer loads: one buffer resource per layer built from the uniform base
// (scalar instructions), the k offset in the scalar offset and the lane's offset in the vector
// offset, so that no per-layer 64-bit address arithmetic runs on the vector ALU (with plain
// pointers the compiler kept a 64-bit per-lane address and added to it before every load).
// Used by the forward (N=1024 L=1024 B=2048: 1.14 ms against 1.17 ms), fwddiff (two tables per
// layer: 2.30 ms against 2.63 ms) and the HVP walk; the adjoint tiles keep plain pointers, which
// measured faster there (N=1024 adjoint 5.40 ms against 5.61 ms with buffer loads).
template <class R>
struct LanePtr {
    const R* base;
    unsigned off;
    __device__ __forceinline__ LanePtr operator+(const size_t n) const { return LanePtr{base + n, off}; }
};

template <class R>
__device__ __forceinline__ LanePtr<R> lane_ptr(const R* base, const int t) {
    return LanePtr<R>{base, (unsigned)t * (unsigned)sizeof(R)};
}

// Raw buffer resource over memory starting at p (gfx9-family word 3: 32-bit data format, no
// swizzle); the range limit is not used as a bounds check here (every offset is in bounds).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t table_rsrc(const void* p) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, 0x7fffffff, 0x00020000);
}

template <class R>
__device__ __forceinline__ R buf_load(const __amdgpu_buffer_rsrc_t rs, const unsigned voff, const int soff) {
    static_assert(sizeof(R) == 16 || sizeof(R) == 8 || sizeof(R) == 4, "row entry size");
    if constexpr (sizeof(R) == 16)
        return __builtin_bit_cast(R, __builtin_amdgcn_raw_buffer_load_b128(rs, voff, soff, 0));
    else if constexpr (sizeof(R) == 8)
        return __builtin_bit_cast(R, __builtin_amdgcn_raw_buffer_load_b64(rs, voff, soff, 0));
    else
        return __builtin_bit_cast(R, __builtin_amdgcn_raw_buffer_load_b32(rs, voff, soff, 0));
}

template <class R, int W>
__device__ __forceinline__ void load_rows(R (&r)[W], const LanePtr<R> p, const int LPV) {
    const __amdgpu_buffer_rsrc_t rs = table_rsrc(p.base);
#pragma unroll
    for (int k = 0; k < W; ++k) r[k] = buf_load<R>(rs, p.off, k * LPV * (int)sizeof(R));
}

// ---- fwdprop --------------------------------------------------------------------------
// Layers [l0, l0 + nl) of the row tables.  The layer loop is unrolled by two with two row
// register sets (rows of layer l+1 are in flight while layer l is applied) and two field
// register sets, so that no result overwrites an input of its own layer.
// Field checkpoints of the lossy adjoint: the field entering every layer l with
// l % every == 0 lives at base + (l / every) * stride (base == nullptr: none).
struct Ckpt {
    void* base;
    int every;
    size_t stride;  // elements
};

// KP: pointer type of the layer-kind array (a constant-address-space pointer where the caller
// can promise it is read-only for the launch, so that its loads are scalar).
template <class T, int W, int V, class KP>
__device__ __forceinline__ void fwd_tile(const int N, const int l0, const int nl, const int B, const int LPV,
                                         const int NP, KP kind,
                                         const typename RowOf<T>::type* __restrict__ rows, const T* u_in, T* u_out,
                                         const int ldu, const long tile, const Ckpt ck = Ckpt{nullptr, 0, 0}) {
    typedef typename RowOf<T>::type R;
    const Lanes q = lanes_of<V>(LPV, tile);
    T ua[V][W], ub[V][W];
    load_tile(ua, u_in, q, N, B, ldu);
    if (nl > 0) {
        const KP kp = kind + l0;
        const LanePtr<R> rp = lane_ptr(rows + (size_t)l0 * NP, q.t);
        R ra[W], rb[W];
        int ka = kp[0];
        load_rows(ra, rp, LPV);
        // Next checkpointed layer (local index) and its slot; INT_MAX when there is none.
        T* const ckb = reinterpret_cast<T*>(ck.base);
        int ck_l = 0x7fffffff, ck_c = 0;
        if (ckb) {
            ck_c = (l0 + ck.every - 1) / ck.every;
            ck_l = ck_c * ck.every - l0;
        }
        int l = 0;
        for (; l + 1 < nl; l += 2) {
            const int kb = kp[l + 1];
            load_rows(rb, rp + (size_t)(l + 1) * NP, LPV);
            if (l == ck_l) {
                store_tile(ua, ckb + (size_t)ck_c * ck.stride, q, N, B, ldu);
                ++ck_c;
                ck_l += ck.every;
            }
            {
                FwdStep<T, W, V> st{ua, ub, ra, q.lane};
                dispatch_kind<W>(ka, st);
            }
            const int ln = (l + 2 < nl) ? l + 2 : nl - 1;
            ka = kp[ln];
            load_rows(ra, rp + (size_t)ln * NP, LPV);
            if (l + 1 == ck_l) {
                store_tile(ub, ckb + (size_t)ck_c * ck.stride, q, N, B, ldu);
                ++ck_c;
                ck_l += ck.every;
            }
            {
                FwdStep<T, W, V> st{ub, ua, rb, q.lane};
                dispatch_kind<W>(kb, st);
            }
        }
        if (l < nl) {
            if (l == ck_l) store_tile(ua, ckb + (size_t)ck_c * ck.stride, q, N, B, ldu);
            FwdStep<T, W, V> st{ua, ub, ra, q.lane};
            dispatch_kind<W>(ka, st);
            store_tile(ub, u_out, q, N, B, ldu);
            return;
        }
    }
    store_tile(ua, u_out, q, N, B, ldu);
}

template <class T, int W, int V>
__global__ void __launch_bounds__(256) k_fwd(int N, int l0, int nl, int B, int LPV, int NP,
                                             const int* __restrict__ kind,
                                             const typename RowOf<T>::type* __restrict__ rows, const T* u_in,
                                             T* u_out, int ldu, Ckpt ck) {
    fwd_tile<T, W, V, const int*>(N, l0, nl, B, LPV, NP, kind, rows, u_in, u_out, ldu, blockIdx.x, ck);
}

// ---- neighbour-mesh forward ------------------------------------------------------------
// NN geometries (layer kinds 0 and 1) whose vectors fill one wave (LPV = 64): one layer per
// loop iteration with the field tile updated in place (one body per shift parity).  Same
// arithmetic per waveguide as fwd_tile (bitwise equal results).
template <class T, int W, int V, bool ODD>
__device__ __forceinline__ void fwd_layer_nn(T (&u)[V][W], const typename RowOf<T>::type (&r)[W]) {
    if constexpr (!ODD) {
#pragma unroll
        for (int k = 0; k < W; k += 2) {
#pragma unroll
            for (int v = 0; v < V; ++v) {
                const T a = u[v][k], b = u[v][k + 1];
                u[v][k] = rapply(r[k], a, b);
                u[v][k + 1] = rapply(r[k + 1], b, a);
            }
        }
    } else {
        T hi[V], lo[V];
#pragma unroll
        for (int v = 0; v < V; ++v) {
            hi[v] = lane_next(u[v][0]);
            lo[v] = lane_prev(u[v][W - 1]);
        }
#pragma unroll
        for (int k = 1; k + 1 < W; k += 2) {
#pragma unroll
            for (int v = 0; v < V; ++v) {
                const T a = u[v][k], b = u[v][k + 1];
                u[v][k] = rapply(r[k], a, b);
                u[v][k + 1] = rapply(r[k + 1], b, a);
            }
        }
#pragma unroll
        for (int v = 0; v < V; ++v) {
            u[v][W - 1] = rapply(r[W - 1], u[v][W - 1], hi[v]);
            u[v][0] = rapply(r[0], u[v][0], lo[v]);
        }
    }
}

// Rows staged through LDS: the workgroup (four waves) copies each layer's row table into a
// ring of three LDS slots once, and every wave reads its rows from there.  A row feeds 8 FMAs
// per vector, so with rows read from L2 by every wave the texture address unit, at 64 B per
// clock and CU, bounds a one-vector tile near half the FMA rate (counters at N=1024 L=1024
// B=2048: TA busy 92 % of the kernel, VALU issue 44 %); through LDS each layer's rows pass
// the TA once per workgroup.  Layer l: read its rows from slot l % 3, request the rows of
// layer l + 2 into registers, apply the layer, write the rows of layer l + 1 (requested one
// layer earlier) to their slot, and one barrier publishes them.  The slot written at layer l
// was last read at layer l - 2, before the barrier of layer l - 1.
template <class T, int W, int V>
__device__ __forceinline__ void fwd_tile_nn(const int N, const int l0, const int nl, const int B, ConstInt* kind,
                                            const typename RowOf<T>::type* __restrict__ rows, const T* u_in,
                                            T* u_out, const int ldu, const long tile, const Ckpt ck,
                                            typename RowOf<T>::type* __restrict__ lrows) {
    typedef typename RowOf<T>::type R;
    constexpr int LPV = 64, NP = 64 * W;
    constexpr int CP = NP / 256;  // entries of a layer each thread copies (workgroups of 256 threads)
    static_assert(NP % 256 == 0, "whole layers per workgroup copy");
    const Lanes q = lanes_of<V>(LPV, tile);
    const int tid = threadIdx.x;
    T u[V][W];
    load_tile(u, u_in, q, N, B, ldu);
    if (nl > 0) {
        const __amdgpu_buffer_rsrc_t rrs = table_rsrc(rows + (size_t)l0 * NP);
        const unsigned coff = (unsigned)tid * (unsigned)sizeof(R);
        R cp[2][CP];  // rows of the next two layers on their way to LDS
        auto fetch = [&](R (&c)[CP], const int l) {
            const int lc = l < nl ? l : nl - 1;  // past the end: the last layer again (in bounds, unused)
#pragma unroll
            for (int e = 0; e < CP; ++e) c[e] = buf_load<R>(rrs, coff, (lc * NP + e * 256) * (int)sizeof(R));
        };
        auto publish = [&](const R (&c)[CP], const int l) {
            R* dst = lrows + (l % 3) * NP + tid;
#pragma unroll
            for (int e = 0; e < CP; ++e) dst[e * 256] = c[e];
        };
        // LDS writes of this thread done, then the workgroup barrier (row requests in flight
        // stay in flight: no wait on the vector memory counter).
        auto lds_barrier = [] { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); };
        fetch(cp[0], 0);
        fetch(cp[1], 1);
        publish(cp[0], 0);
        lds_barrier();
        T* const ckb = reinterpret_cast<T*>(ck.base);
        int ck_l = 0x7fffffff, ck_c = 0;
        if (ckb) {
            ck_c = (l0 + ck.every - 1) / ck.every;
            ck_l = ck_c * ck.every - l0;
        }
        int kd = kind[l0];
        // Two layers per iteration: the register sets of the requested rows alternate.
        auto layer = [&](const int l, R (&cnext)[CP], R (&cafter)[CP]) {
            const int kn = kind[l0 + (l + 1 < nl ? l + 1 : l)];
            const R* cur = lrows + (l % 3) * NP + q.t;
            R r[W];
#pragma unroll
            for (int k = 0; k < W; ++k) r[k] = cur[k * LPV];
            fetch(cafter, l + 2);
            if (l == ck_l) {
                store_tile(u, ckb + (size_t)ck_c * ck.stride, q, N, B, ldu);
                ++ck_c;
                ck_l += ck.every;
            }
            if (kd & 1)
                fwd_layer_nn<T, W, V, true>(u, r);
            else
                fwd_layer_nn<T, W, V, false>(u, r);
            publish(cnext, l + 1);
            lds_barrier();
            kd = kn;
        };
        int l = 0;
        for (; l + 1 < nl; l += 2) {
            layer(l, cp[1], cp[0]);
            layer(l + 1, cp[0], cp[1]);
        }
        if (l < nl) layer(l, cp[1], cp[0]);
    }
    store_tile(u, u_out, q, N, B, ldu);
}

template <class T, int W, int V>
__global__ void __launch_bounds__(256) k_fwd_nn(int N, int l0, int nl, int B, const int* __restrict__ kind,
                                                const typename RowOf<T>::type* __restrict__ rows, const T* u_in,
                                                T* u_out, int ldu, Ckpt ck) {
    typedef typename RowOf<T>::type R;
    __shared__ R lrows[3 * 64 * W];
    fwd_tile_nn<T, W, V>(N, l0, nl, B, (ConstInt*)kind, rows, u_in, u_out, ldu, blockIdx.x, ck, lrows);
}

// ---- fwddiff --------------------------------------------------------------------------
// NN-specialised forward-mode walk of the widest tile (W = 16, one vector per wave): u and du
// updated in place per layer, the layer parity a compile-time body, both row tables staged
// through LDS as the forward's (k_fwd_nn): a three-slot ring of (rows, tangent rows) per
// layer, requested two layers ahead by the NW waves of the workgroup, each crossing's rows read
// from LDS where it is applied.  Each wave reading 32 KiB of rows per layer through the address
// unit bounds the general fwddiff tile.
template <class T, int W, bool ODD>
__device__ __forceinline__ void fwddiff_layer_nn(T (&u)[W], T (&du)[W], const typename RowOf<T>::type* lr,
                                                 const typename RowOf<T>::type* ld) {
    typedef typename RowOf<T>::type R;
    constexpr int H = W / 2;
    auto pair = [&](const int k, const int j) __attribute__((always_inline)) {
        const R rk = lr[k * 64], rj = lr[j * 64], ek = ld[k * 64], ej = ld[j * 64];
        const T a = u[k], b = u[j], da = du[k], db = du[j];
        u[k] = rapply(rk, a, b);
        u[j] = rapply(rj, b, a);
        du[k] = rapply_d(rk, ek, a, b, da, db);
        du[j] = rapply_d(rj, ej, b, a, db, da);
    };
    if constexpr (!ODD) {
#pragma unroll
        for (int p = 0; p < H; ++p) pair(2 * p, 2 * p + 1);
    } else {
        const T hu = lane_next(u[0]), hd = lane_next(du[0]), lu = lane_prev(u[W - 1]), lo = lane_prev(du[W - 1]);
#pragma unroll
        for (int p = 0; p + 1 < H; ++p) pair(2 * p + 1, 2 * p + 2);
        const R rh = lr[(W - 1) * 64], eh = ld[(W - 1) * 64], rl = lr[0], el = ld[0];
        const T a = u[W - 1], b = u[0], da = du[W - 1], db = du[0];
        u[W - 1] = rapply(rh, a, hu);
        du[W - 1] = rapply_d(rh, eh, a, hu, da, hd);
        u[0] = rapply(rl, b, lu);
        du[0] = rapply_d(rl, el, b, lu, db, lo);
    }
}

template <class T, int W, int NW>
__device__ __forceinline__ void fwddiff_tile_nn(const int N, const int nl, const int B, ConstInt* kind,
                                                const typename RowOf<T>::type* __restrict__ rows,
                                                const typename RowOf<T>::type* __restrict__ drows, const T* u_in,
                                                const T* du_in, T* u_out, T* du_out, const int ldu, const long tile,
                                                typename RowOf<T>::type* __restrict__ lrows) {
    typedef typename RowOf<T>::type R;
    constexpr int LPV = 64, NP = 64 * W, NTHR = 64 * NW;
    constexpr int CP = NP / NTHR;  // entries of a layer's table each thread copies
    static_assert(NP % NTHR == 0, "whole layers per workgroup copy");
    const Lanes q = lanes_of<1>(LPV, tile);
    const int tid = threadIdx.x;
    T u[1][W], du[1][W];
    load_tile(u, u_in, q, N, B, ldu);
    if (du_in != nullptr) {
        load_tile(du, du_in, q, N, B, ldu);
    } else {
#pragma unroll
        for (int k = 0; k < W; ++k) du[0][k] = fzero(T());
    }
    if (nl > 0) {
        const __amdgpu_buffer_rsrc_t rrs = table_rsrc(rows), drs = table_rsrc(drows);
        const unsigned coff = (unsigned)tid * (unsigned)sizeof(R);
        R cp[2][2][CP];  // [register set][table][entry]: rows of the next two layers on their way to LDS
        auto fetch = [&](R (&c)[2][CP], const int l) {
            const int lc = l < nl ? l : nl - 1;  // past the end: the last layer again (in bounds, unused)
#pragma unroll
            for (int e = 0; e < CP; ++e) {
                c[0][e] = buf_load<R>(rrs, coff, (lc * NP + e * NTHR) * (int)sizeof(R));
                c[1][e] = buf_load<R>(drs, coff, (lc * NP + e * NTHR) * (int)sizeof(R));
            }
        };
        auto publish = [&](const R (&c)[2][CP], const int l) {
            R* dst = lrows + (l % 3) * 2 * NP + tid;
#pragma unroll
            for (int e = 0; e < CP; ++e) {
                dst[e * NTHR] = c[0][e];
                dst[NP + e * NTHR] = c[1][e];
            }
        };
        auto lds_barrier = [] { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); };
        fetch(cp[0], 0);
        fetch(cp[1], 1);
        publish(cp[0], 0);
        lds_barrier();
        int kd = kind[0];
        auto layer = [&](const int l, R (&cnext)[2][CP], R (&cafter)[2][CP]) {
            const int kn = kind[l + 1 < nl ? l + 1 : l];
            const R* cur = lrows + (l % 3) * 2 * NP + q.t;
            fetch(cafter, l + 2);
            if (kd & 1)
                fwddiff_layer_nn<T, W, true>(u[0], du[0], cur, cur + NP);
            else
                fwddiff_layer_nn<T, W, false>(u[0], du[0], cur, cur + NP);
            publish(cnext, l + 1);
            lds_barrier();
            kd = kn;
        };
        int l = 0;
        for (; l + 1 < nl; l += 2) {
            layer(l, cp[1], cp[0]);
            layer(l + 1, cp[0], cp[1]);
        }
        if (l < nl) layer(l, cp[1], cp[0]);
    }
    store_tile(u, u_out, q, N, B, ldu);
    if (du_out != nullptr) store_tile(du, du_out, q, N, B, ldu);
}

constexpr int kFwdDiffNNWaves = 8;

template <class T, int W>
__global__ void __launch_bounds__(64 * kFwdDiffNNWaves) k_fwddiff_nn(int N, int L, int B, const int* __restrict__ kind,
                                                                   const typename RowOf<T>::type* __restrict__ rows,
                                                                   const typename RowOf<T>::type* __restrict__ drows,
                                                                   const T* u_in, const T* du_in, T* u_out, T* du_out,
                                                                   int ldu) {
    typedef typename RowOf<T>::type R;
    extern __shared__ unsigned char smem_raw[];
    fwddiff_tile_nn<T, W, kFwdDiffNNWaves>(N, L, B, (ConstInt*)kind, rows, drows, u_in, du_in, u_out, du_out, ldu,
                                           blockIdx.x, reinterpret_cast<R*>(smem_raw));
}

template <class T, int W, int V>
__global__ void __launch_bounds__(256) k_fwddiff(int N, int L, int B, int LPV, int NP, const int* __restrict__ kind,
                                                 const typename RowOf<T>::type* __restrict__ rows,
                                                 const typename RowOf<T>::type* __restrict__ drows, const T* u_in,
                                                 const T* du_in, T* u_out, T* du_out, int ldu) {
    typedef typename RowOf<T>::type R;
    const Lanes q = lanes_of<V>(LPV, blockIdx.x);
    T u[V][W], du[V][W], ub[V][W], dub[V][W];
    load_tile(u, u_in, q, N, B, ldu);
    if (du_in != nullptr) {
        load_tile(du, du_in, q, N, B, ldu);
    } else {
#pragma unroll
        for (int v = 0; v < V; ++v)
#pragma unroll
            for (int k = 0; k < W; ++k) du[v][k] = fzero(T());
    }
    if (L > 0) {
        const LanePtr<R> rp = lane_ptr(rows, q.t);
        const LanePtr<R> dq = lane_ptr(drows, q.t);
        R ra[W], da[W], rb[W], db[W];
        int ka = kind[0];
        load_rows(ra, rp, LPV);
        load_rows(da, dq, LPV);
        int l = 0;
        for (; l + 1 < L; l += 2) {
            const int kb = kind[l + 1];
            load_rows(rb, rp + (size_t)(l + 1) * NP, LPV);
            load_rows(db, dq + (size_t)(l + 1) * NP, LPV);
            {
                FwdDiffStep<T, W, V> st{u, du, ub, dub, ra, da, q.lane};
                dispatch_kind<W>(ka, st);
            }
            const int ln = (l + 2 < L) ? l + 2 : L - 1;
            ka = kind[ln];
            load_rows(ra, rp + (size_t)ln * NP, LPV);
            load_rows(da, dq + (size_t)ln * NP, LPV);
            {
                FwdDiffStep<T, W, V> st{ub, dub, u, du, rb, db, q.lane};
                dispatch_kind<W>(kb, st);
            }
        }
        if (l < L) {
            FwdDiffStep<T, W, V> st{u, du, ub, dub, ra, da, q.lane};
            dispatch_kind<W>(ka, st);
            store_tile(ub, u_out, q, N, B, ldu);
            if (du_out != nullptr) store_tile(dub, du_out, q, N, B, ldu);
            return;
        }
    }
    store_tile(u, u_out, q, N, B, ldu);
    if (du_out != nullptr) store_tile(du, du_out, q, N, B, ldu);
}

// ---- backdiff -------------------------------------------------------------------------
// Workgroup = nw waves.  Per layer every lane accumulates its dJdT rows over its V
// vectors in registers; lane groups are folded with cross-lane adds; the wave writes the
// rows (unitary crossings: one K-form entry per crossing slot) to an LDS stage of C layers.
// At the end of each stage the workgroup sums the waves in a fixed order, converts to dp
// with the K-form or dT/dp table and writes (or, from its second vector tile on, adds) a
// per-workgroup partial.  The dp reduction kernels then
// sum the partials in a fixed order: the gradient is bitwise reproducible (the reference
// uses shared and global float atomics, gpu/crossing.cu:179-180, 295-296).
// Layers [l0, l0 + nl) are walked from the last to the first; checkpointed adjoints run
// one launch per segment.
template <class T>
__device__ __forceinline__ float re_dot(const T a, const T b);
template <>
__device__ __forceinline__ float re_dot<float2>(const float2 a, const float2 b) {
    return fmaf(a.x, b.x, -a.y * b.y);
}
template <>
__device__ __forceinline__ float re_dot<float>(const float a, const float b) {
    return a * b;
}

// K-form stage entry of one crossing, i its upper and j its lower waveguide, from their rows
// (A = Im part in .x, B = conj(g') x_partner): (D_i, D_j, B_i - conj(B_j)) for complex fields,
// (B_i - B_j, 0) for real fields.
__device__ __forceinline__ float4 kquad(const float2 ai, const float2 aj, const float2 bi, const float2 bj) {
    return make_float4(ai.x, aj.x, bi.x - bj.x, bi.y + bj.y);
}
__device__ __forceinline__ float2 kquad(const float, const float, const float bi, const float bj) {
    return make_float2(bi - bj, 0.f);
}
// Per-lane choice between two entries, component by component (a select of the vector
// structs is lowered through a stack temporary).
__device__ __forceinline__ float4 esel(const bool c, const float4 x, const float4 y) {
    return make_float4(c ? x.x : y.x, c ? x.y : y.y, c ? x.z : y.z, c ? x.w : y.w);
}
__device__ __forceinline__ float2 esel(const bool c, const float2 x, const float2 y) {
    return make_float2(c ? x.x : y.x, c ? x.y : y.y);
}
// Cross-lane reads of a K-form A row (.x only) and of a B row.
__device__ __forceinline__ float2 a_next(const float2 a) { return make_float2(lane_next(a.x), 0.f); }
__device__ __forceinline__ float a_next(const float) { return 0.f; }
__device__ __forceinline__ float2 a_read(const float2 a, const int src) { return make_float2(lane_read(a.x, src), 0.f); }
__device__ __forceinline__ float a_read(const float, const int) { return 0.f; }

// Writes the K-form stage entries of one layer, slot-major: entry m is the crossing of slot m
// (slot numbering of the prologue), ws the wave's stage layer.  In-lane pairings (NN layers,
// FFT strides below W, identity layers) give lane t the slots t*W/2 .. t*W/2 + W/2 - 1; odd NN
// layers take the last pair's lower waveguide from lane t+1; FFT strides of W and more pair
// lane t with t ^ xm, and each of the two lanes writes half of the upper lane's W slots.
// Every slot below NP/2 is written, so the conversion never reads stale LDS.  The lane
// exchanges run with every lane active; only the first lane group writes.
template <class T, int W>
__device__ __forceinline__ void kstage_write(const int kd, const T (&a)[W], const T (&b)[W],
                                             typename RowOf<T>::type* ws, const Lanes& q) {
    constexpr int H = W / 2;
    constexpr int LW = Log2<W>::value;
    const bool wr = q.grp == 0;
    typename RowOf<T>::type* wl = ws + q.t * H;
    if (kd == 1) {
        const T na = a_next(a[0]), nb = lane_next(b[0]);
        if (wr) {
#pragma unroll
            for (int p = 0; p + 1 < H; ++p) wl[p] = kquad(a[2 * p + 1], a[2 * p + 2], b[2 * p + 1], b[2 * p + 2]);
            wl[H - 1] = kquad(a[W - 1], na, b[W - 1], nb);
        }
        return;
    }
    const int st = (kd == 0 || kd == KIND_IDENTITY) ? 0 : kd - 2;
    if (st < LW) {
        if (!wr) return;
        auto inlane = [&](auto sc) {
            constexpr int S = decltype(sc)::value;
#pragma unroll
            for (int k = 0; k < W; ++k)
                if ((k & S) == 0) wl[(k / (2 * S)) * S + (k % S)] = kquad(a[k], a[k + S], b[k], b[k + S]);
        };
        if (st == 0) inlane(std::integral_constant<int, 1>());
        if constexpr (W > 2) {
            if (st == 1) inlane(std::integral_constant<int, 2>());
        }
        if constexpr (W > 4) {
            if (st == 2) inlane(std::integral_constant<int, 4>());
        }
        if constexpr (W > 8) {
            if (st == 3) inlane(std::integral_constant<int, 8>());
        }
        return;
    }
    const int lx = st - LW, xm = 1 << lx;
    const bool top = (q.t & xm) == 0;
    const int src = q.lane ^ xm;
    // Upper lane tt of the pair: its W slots start at ((tt / 2xm) * xm + tt % xm) * W (xm is a
    // power of two: shifts and masks).
    const int tt = top ? q.t : (q.t ^ xm);
    typename RowOf<T>::type* wt = ws + ((((tt >> (lx + 1)) << lx) | (tt & (xm - 1))) << LW) + (top ? 0 : H);
    if constexpr (H == 1) {
        // W = 2: the upper lane writes crossing 0 (its waveguide 0, the lower lane's 0), the
        // lower lane crossing 1 (the upper lane's waveguide 1, its own 1).  Both waveguides
        // are exchanged: a lane-dependent pick between the two array elements would be
        // lowered to a stack access.
        const T ra0 = a_read(a[0], src), rb0 = lane_read(b[0], src);
        const T ra1 = a_read(a[1], src), rb1 = lane_read(b[1], src);
        const auto e_up = kquad(a[0], ra0, b[0], rb0), e_lo = kquad(ra1, a[1], rb1, b[1]);
        if (wr) wt[0] = esel(top, e_up, e_lo);
    } else {
#pragma unroll
        for (int k = 0; k < H; ++k) {
            // The upper lane keeps waveguides k < H and needs the lower lane's; the lower lane
            // keeps k >= H and needs the upper lane's.
            const T ra = a_read(top ? a[k + H] : a[k], src);
            const T rb = lane_read(top ? b[k + H] : b[k], src);
            const auto e = esel(top, kquad(a[k], ra, b[k], rb), kquad(ra, a[k + H], rb, b[k + H]));
            if (wr) wt[k] = e;
        }
    }
}

// Lane-group fold of a K-form A row: only .x carries data (real fields carry none).
__device__ __forceinline__ float2 fold_im(const float2 a, const int src) {
    return make_float2(a.x + lane_read(a.x, src), a.y);
}
__device__ __forceinline__ float fold_im(const float a, const int) { return a; }

template <class T, int W, int V, bool UNI, bool SC = false>
__device__ __forceinline__ void bwd_layer(const int kd, const T (&u)[V][W], const T (&g)[V][W], T (&uo)[V][W],
                                          T (&go)[V][W], const typename RowOf<T>::type (&ri)[W],
                                          const typename RowOf<T>::type (&ra)[W], typename RowOf<T>::type* stage,
                                          const Lanes& q, const int LPV, const int C, const int lc,
                                          const float* scl = nullptr) {
    T accA[W], accB[W];
#pragma unroll
    for (int k = 0; k < W; ++k) {
        accA[k] = fzero(T());
        accB[k] = fzero(T());
    }
    // Unitary crossings take the K-form rows (A_k.x = Im part, A_k.y = 0; no A for real fields).
    BwdStep<T, W, V, UNI, SC> st{u, g, uo, go, ri, ra, accA, accB, q.lane, scl};
    dispatch_kind<W>(kd, st);
    // Fold lane groups (groups hold the same waveguides of different vectors).
    for (int off = LPV; off < 64; off <<= 1) {
#pragma unroll
        for (int k = 0; k < W; ++k) {
            if constexpr (!UNI)
                accA[k] = fadd(accA[k], lane_read(accA[k], q.lane ^ off));
            else
                accA[k] = fold_im(accA[k], q.lane ^ off);
            accB[k] = fadd(accB[k], lane_read(accB[k], q.lane ^ off));
        }
    }
    if constexpr (UNI) {
        // K-form: one entry per crossing slot, NP/2 per wave and layer.
        kstage_write<T, W>(kd, accA, accB,
                           stage + (size_t)(__builtin_amdgcn_readfirstlane(q.wave) * C + lc) * (W / 2) * LPV, q);
    } else if (q.grp == 0) {
        // This wave's rows of stage layer lc (a wave-uniform base), then k * LPV + lane.
        typename RowOf<T>::type* ws =
            stage + (size_t)(__builtin_amdgcn_readfirstlane(q.wave) * C + lc) * W * LPV + q.t;
#pragma unroll
        for (int k = 0; k < W; ++k) ws[k * LPV] = pack_row(accA[k], accB[k]);
    }
}

// Stage-end conversion of one slot: sum the waves' dJdT rows of slot m (layer kind kd,
// wave-uniform) from the stage layer s0 and add its dp contribution to pout[m*np ...].
__device__ __forceinline__ int s_clamp(const int last, const int s) { return s < last ? last - s : 0; }

// Most waves that share one LDS stage: every kernel that converts a stage (backdiff, the
// grouped kernel, the Hessian-vector product) runs workgroups of at most 256 threads.
constexpr int kStageWaves = 4;

// K-form dp of one parameter from the summed stage entry e = (Im Q00, Im Q11, Q01 - conj(Q10))
// of a slot: -a Im Q00 - d Im Q11 + Re(c (Q01 - conj(Q10))); real fields: c (Q01 - Q10).
__device__ __forceinline__ float kdot(const float4 e, const float4 c) {
    return fmaf(-c.w, e.w, fmaf(c.z, e.z, fmaf(-c.y, e.y, -c.x * e.x)));
}
__device__ __forceinline__ float kdot(const float2 e, const float c) { return c * e.x; }

// dp of each parameter of a slot from its summed rows i, j: dJdT . dT/dp
// (dJdT00 = A_i, dJdT01 = B_i, dJdT11 = A_j, dJdT10 = B_j).
template <class T, int NPAR, class R, class CT>
__device__ __forceinline__ void slot_dp(const R si, const R sj, const CT* dd, float (&val)[NPAR]) {
#pragma unroll
    for (int k = 0; k < NPAR; ++k) {
        val[k] = re_dot(row_a(si), dd[k * 4 + 0]);
        val[k] += re_dot(row_b(si), dd[k * 4 + 1]);
        val[k] += re_dot(row_b(sj), dd[k * 4 + 2]);
        val[k] += re_dot(row_a(sj), dd[k * 4 + 3]);
    }
}

// Stores (first vector tile) or adds a slot's dp values.
template <int NPAR>
__device__ __forceinline__ void put_dp(float* outp, const float (&val)[NPAR], const bool first_tile) {
    if constexpr (NPAR == 2) {
        float2* o2 = reinterpret_cast<float2*>(outp);
        if (first_tile) {
            *o2 = make_float2(val[0], val[1]);
        } else {
            const float2 old = *o2;
            *o2 = make_float2(old.x + val[0], old.y + val[1]);
        }
    } else {
        outp[0] = first_tile ? val[0] : outp[0] + val[0];
    }
}

// K-form conversion of slot m: sum the waves' stage entries (slot-major layers, wave stride
// wstride) in wave order and write the slot's dp.  Inactive slots have zero coefficients.
template <class T, int NPAR>
__device__ __forceinline__ void convert_slot_k(const typename RowOf<T>::type* s0, const int wstride, const int nw,
                                               float* pout, const int m, const typename KOf<T>::type (&kc)[NPAR],
                                               const bool first_tile) {
    typedef typename RowOf<T>::type R;
    R e_[kStageWaves];
#pragma unroll
    for (int wv = 0; wv < kStageWaves; ++wv)
        if (wv < nw) e_[wv] = s0[wv * wstride + m];
    R e = e_[0];
#pragma unroll
    for (int wv = 1; wv < kStageWaves; ++wv)
        if (wv < nw) e = row_add(e, e_[wv]);
    float val[NPAR];
#pragma unroll
    for (int k = 0; k < NPAR; ++k) val[k] = kdot(e, kc[k]);
    put_dp<NPAR>(pout + m * NPAR, val, first_tile);
}

template <class T, int W, int NPAR>
__device__ __forceinline__ void convert_slot(const typename RowOf<T>::type* s0, const int wstride, const int nw,
                                             const int LPV, const int NP, float* pout, const int kd, const int m,
                                             const T (&dd)[4 * NPAR], const bool first_tile) {
    typedef typename RowOf<T>::type R;
    int i, j;
    if (kd <= 1) {
        i = 2 * m + kd;
        j = i + 1;
    } else if (kd != KIND_IDENTITY) {
        const int st = kd - 2;
        i = ((m >> st) << (st + 1)) + (m & ((1 << st) - 1));
        j = i + (1 << st);
    } else {
        i = -1;
        j = 0;
    }
    float* outp = pout + m * NPAR;
    if (i < 0 || j >= NP) {
        if (first_tile)
#pragma unroll
            for (int k = 0; k < NPAR; ++k) outp[k] = 0.f;
        return;
    }
    const int oi = (i % W) * LPV + i / W, oj = (j % W) * LPV + j / W;
    // All stage reads are issued before the first sum; the waves are summed in index order.
    R ri_[kStageWaves], rj_[kStageWaves];
#pragma unroll
    for (int wv = 0; wv < kStageWaves; ++wv)
        if (wv < nw) {
            ri_[wv] = s0[wv * wstride + oi];
            rj_[wv] = s0[wv * wstride + oj];
        }
    R si = ri_[0], sj = rj_[0];
#pragma unroll
    for (int wv = 1; wv < kStageWaves; ++wv)
        if (wv < nw) {
            si = row_add(si, ri_[wv]);
            sj = row_add(sj, rj_[wv]);
        }
    float val[NPAR];
    slot_dp<T, NPAR>(si, sj, dd, val);
    put_dp<NPAR>(outp, val, first_tile);
}

// Workgroups of the unpipelined adjoint stage hold at most four waves (eight-wave workgroups
// exist only with the pipelined stage, kPipeWaves).
__host__ __device__ constexpr int bwd_max_waves(int, int, bool) { return 4; }
static_assert(bwd_max_waves(0, 0, true) <= kStageWaves, "stage conversion sums at most kStageWaves waves");

// Number of phase parameters per crossing, fixed by the field type: real fields come only
// from orthogonal crossings (one parameter), complex fields from the MZI family (two).
template <class T>
struct NPhase {
    static constexpr int value = 2;
};
template <>
struct NPhase<float> {
    static constexpr int value = 1;
};

// Row pipeline depth of backdiff: four register sets (rows requested two layers ahead) for
// narrow tiles, two (one layer ahead) when the tile's registers do not allow more.
template <int W>
struct BwdDepth {
    static constexpr int value = W <= 4 ? 4 : 2;
};

// Per-waveguide factors of one layer (same k-major layout as the row tables).
template <int W>
__device__ __forceinline__ void load_scl(float (&s)[W], const float* p, const int LPV) {
#pragma unroll
    for (int k = 0; k < W; ++k) s[k] = p[(size_t)k * LPV];
}

template <class T, int W, int V, bool UNI, class KP, bool SC = false>
__device__ __forceinline__ void bwd_tile(const int N, const int l0, const int nl, const int B, const int LPV,
                                         const int NP, const int C, KP kind,
                                         const typename RowOf<T>::type* __restrict__ rows_inv,
                                         const typename RowOf<T>::type* __restrict__ rows_adj,
                                         const T* __restrict__ dtdp, float* __restrict__ my_partial, const T* u_out,
                                         const T* g_out, const T* target, T* u_in, T* g_in, const int ldu,
                                         const long tile, const bool first,
                                         typename RowOf<T>::type* __restrict__ stage,
                                         const Ckpt ck = Ckpt{nullptr, 0, 0}, const float* __restrict__ scales = nullptr) {
    typedef typename RowOf<T>::type R;
    constexpr int np = NPhase<T>::value;
    // Conversion coefficients: K-form for unitary crossings, dT/dp otherwise.
    typedef typename CoefOf<T, UNI>::type CT;
    constexpr int NC = np * CoefOf<T, UNI>::per_param;
    const CT* ctab = reinterpret_cast<const CT*>(dtdp);
    constexpr int D = BwdDepth<W>::value;  // row register sets
    constexpr int PD = D / 2;              // prefetch distance in layers
    const int nw = blockDim.x >> 6;
    const int NPh = NP / 2;
    const int lane = threadIdx.x & 63;
    const int wave_u = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave index as a scalar
    const int lrows = UNI ? NPh : NP;  // stage entries per wave and layer (K-form: one per slot)
    const int wstride = C * lrows;
    const KP kp = kind + l0;
    const int last = nl - 1;
    // Lean form for the wide non-unitary one-vector tiles (two row tables of W rows per layer):
    // no dT/dp prefetch and the adjoint rows loaded just before their layer, which keeps the
    // kernel within the registers of two waves per SIMD (measured at N=512, L=512, B=8192:
    // 7.2 ms against 11.6 ms for the full form at one wave per SIMD; the two-vector tile in
    // its full form, one wave per SIMD, is faster still at 6.8 ms).
    constexpr bool kLean = (W >= 8) && !UNI && V == 1;
    // Lean unitary form for the widest tile (W = 16): its rows are loaded right before their
    // layer instead of one layer ahead, so one row set is live instead of two.
    constexpr bool kLeanU = UNI && W >= 16;
    constexpr int PF = (kLean || kLeanU) ? 0 : 2;  // stage-end slots per lane whose dT/dp is prefetched
    constexpr int DA = kLean ? 1 : D;  // register sets of adjoint rows
    const Lanes q = lanes_of<V>(LPV, tile);
    T ua[V][W], ga[V][W], ub[V][W], gb[V][W];
    load_tile(ua, u_out, q, N, B, ldu);
    if (target != nullptr) {
        T y[V][W];
        load_tile(y, target, q, N, B, ldu);
#pragma unroll
        for (int v = 0; v < V; ++v)
#pragma unroll
            for (int k = 0; k < W; ++k) ga[v][k] = fsub2(ua[v][k], y[v][k]);
    } else {
        load_tile(ga, g_out, q, N, B, ldu);
    }
    // Local layers are visited in the order s = 0, 1, ..., as l = last - s.
    const R* rip = rows_inv + (size_t)l0 * NP + q.t;
    const R* rap = rows_adj + (size_t)l0 * NP + q.t;
    R ri[D][W], ra[DA][W];
    constexpr int SD = SC ? D : 1;  // register sets of per-waveguide adjoint factors
    float sc[SD][SC ? W : 1];
    const float* scp = SC ? scales + (size_t)l0 * NP + q.t : nullptr;
    int kk[D];
    auto lrow = [&](const int s) { return (size_t)(s < last ? last - s : 0) * NP; };
#pragma unroll
    for (int d = 0; d < PD; ++d) {
        if (nl > 0) {
            kk[d] = kp[s_clamp(last, d)];
            if constexpr (!kLeanU) load_rows(ri[d], rip + lrow(d), LPV);
            if constexpr (SC && !kLeanU) load_scl(sc[d], scp + lrow(d), LPV);
            if constexpr (!UNI && !kLean) load_rows(ra[d], rap + lrow(d), LPV);
        }
    }
#define MP_BWD_LAYER(KD, U, G, UO, GO, RI, RA, LC, SCL)                                             \
    {                                                                                               \
    if constexpr (UNI)                                                                          \
        bwd_layer<T, W, V, UNI, SC>(KD, U, G, UO, GO, RI, RI, stage, q, LPV, C, LC, SCL);       \
    else                                                                                        \
        bwd_layer<T, W, V, UNI>(KD, U, G, UO, GO, RI, RA, stage, q, LPV, C, LC);                \
    }
    // Adjoint rows of local layer SL into set AD (lean form: loaded right before the layer).
#define MP_ADJ_ROWS(AD, SL)                                                                         \
    if constexpr (!UNI && kLean) load_rows(ra[AD], rap + lrow(SL), LPV);
    // Rows (and scaled-unitary factors) of local layer SL into set RD (lean unitary form).
#define MP_LEAN_U_ROWS(RD, SL)                                                                      \
    if constexpr (kLeanU) {                                                                         \
        load_rows(ri[RD], rip + lrow(SL), LPV);                                                     \
        if constexpr (SC) load_scl(sc[(RD) % SD], scp + lrow(SL), LPV);                             \
    }
    for (int c0 = 0; c0 < nl; c0 += C) {
        const int cl = (C < nl - c0) ? C : nl - c0;
        if (ck.base != nullptr && c0 > 0) {
            // The field entering layer e is stored: replace the reconstruction by it (the host
            // only passes checkpoints whose layers fall on stage starts).
            const int e = l0 + nl - c0;
            if (e % ck.every == 0)
                load_tile(ua, reinterpret_cast<const T*>(ck.base) + (size_t)(e / ck.every) * ck.stride, q, N, B, ldu);
        }
        // Prefetch the coefficients this wave converts first at the end of the stage.
        CT pre[PF > 0 ? PF : 1][NC];
        if (PF > 0 && wave_u < cl) {
            const CT* src = ctab + (size_t)(l0 + last - (c0 + wave_u)) * NPh * NC + lane;
#pragma unroll
            for (int r = 0; r < PF; ++r)
                if (lane + 64 * r < NPh) {
#pragma unroll
                    for (int e = 0; e < NC; ++e) pre[r][e] = src[e * NPh + 64 * r];
                }
        }
        int lc = 0;
        for (; lc + D <= cl; lc += D) {
            const int s = c0 + lc;
#pragma unroll
            for (int d = 0; d < D; ++d) {
                const int dn = (d + PD) % D;  // register set receiving layer s + d + PD
                kk[dn] = kp[s_clamp(last, s + d + PD)];
                if constexpr (!kLeanU) load_rows(ri[dn], rip + lrow(s + d + PD), LPV);
                if constexpr (SC && !kLeanU) load_scl(sc[dn % SD], scp + lrow(s + d + PD), LPV);
                if constexpr (!UNI && !kLean) load_rows(ra[dn % DA], rap + lrow(s + d + PD), LPV);
                MP_ADJ_ROWS(0, s + d)
                MP_LEAN_U_ROWS(d, s + d)
                if ((d & 1) == 0)
                    MP_BWD_LAYER(kk[d], ua, ga, ub, gb, ri[d], ra[d % DA], lc + d, sc[d % SD])
                else
                    MP_BWD_LAYER(kk[d], ub, gb, ua, ga, ri[d], ra[d % DA], lc + d, sc[d % SD])
            }
        }
        // Remainder of the last stage (cl - lc < D layers): register sets 0.. hold them.
        int left = cl - lc;
#pragma unroll
        for (int d = 0; d < D - 1; ++d) {
            if (d < left) {
                if (d + PD < D) {
                    const int dn = (d + PD) % D;
                    kk[dn] = kp[s_clamp(last, c0 + lc + d + PD)];
                    if constexpr (!kLeanU) load_rows(ri[dn], rip + lrow(c0 + lc + d + PD), LPV);
                    if constexpr (SC && !kLeanU) load_scl(sc[dn % SD], scp + lrow(c0 + lc + d + PD), LPV);
                    if constexpr (!UNI && !kLean) load_rows(ra[dn % DA], rap + lrow(c0 + lc + d + PD), LPV);
                }
                MP_ADJ_ROWS(0, c0 + lc + d)
                MP_LEAN_U_ROWS(d, c0 + lc + d)
                if ((d & 1) == 0)
                    MP_BWD_LAYER(kk[d], ua, ga, ub, gb, ri[d], ra[d % DA], lc + d, sc[d % SD])
                else
                    MP_BWD_LAYER(kk[d], ub, gb, ua, ga, ri[d], ra[d % DA], lc + d, sc[d % SD])
            }
        }
        if (left & 1) {
            copy_state(ub, ua);
            copy_state(gb, ga);
        }
        __syncthreads();
        // dJdT -> dp: wave w converts stage layers w, w + nw, ... (the first one prefetched).
        for (int lcc = wave_u; lcc < cl; lcc += nw) {
            const int la = l0 + last - (c0 + lcc);
            const R* s0 = stage + lcc * lrows;
            float* pout = my_partial + (size_t)la * NPh * np;
            const CT* src = ctab + (size_t)la * NPh * NC;
            auto conv = [&](const int m, const CT (&dd)[NC]) {
                if constexpr (UNI)
                    convert_slot_k<T, np>(s0, wstride, nw, pout, m, dd, first);
                else
                    convert_slot<T, W, np>(s0, wstride, nw, LPV, NP, pout, kind[la], m, dd, first);
            };
            int m0 = lane;
            if (lcc == wave_u) {
#pragma unroll
                for (int r = 0; r < PF; ++r)
                    if (lane + 64 * r < NPh) conv(lane + 64 * r, pre[r]);
                m0 = lane + 64 * PF;
            }
            for (int m = m0; m < NPh; m += 64) {
                CT dd[NC];
#pragma unroll
                for (int e = 0; e < NC; ++e) dd[e] = src[e * NPh + m];
                conv(m, dd);
            }
        }
        __syncthreads();
    }
#undef MP_BWD_LAYER
#undef MP_ADJ_ROWS
#undef MP_LEAN_U_ROWS
    if (u_in != nullptr) store_tile(ua, u_in, q, N, B, ldu);
    if (g_in != nullptr) store_tile(ga, g_in, q, N, B, ldu);
}

// ---- software-pipelined adjoint tile ----------------------------------------------------
// Same arithmetic and the same fixed summation orders as bwd_tile, with the LDS stage
// double-buffered (two buffers of C layers): the dJdT -> dp conversion of stage s runs while
// stage s + 1 propagates.  Its stage reads and table loads are issued just before the first
// layer of stage s + 1 and consumed right after it, so their latency hides under that
// layer's arithmetic instead of stalling every wave of the workgroup between two barriers;
// one barrier per stage remains.  Conversion items (layer of the stage, slot) are spread over
// all threads; the host uses this form when a wave's items lie in one layer (NP/2 a multiple
// of 64) and two buffers of at least two layers fit the LDS budget.
// The pipelined stage serves workgroups of up to eight waves (one per CU at 512 threads);
// NW is the most waves a kernel instance sums (its item holds one row pair per wave).
constexpr int kPipeWaves = 8;

template <class T, int NW, bool KF>
struct ConvItem {
    typename RowOf<T>::type a[NW], b[KF ? 1 : NW];  // K-form: one slot-major entry per wave
    typename CoefOf<T, KF>::type dd[NPhase<T>::value * CoefOf<T, KF>::per_param];
    float* out;
    int state;  // 0: nothing, 1: inactive slot (zero), 2: convert
};

template <class T, int W, int NW, bool KF, class KP>
__device__ __forceinline__ void conv_issue(ConvItem<T, NW, KF>& it, const typename RowOf<T>::type* s0, const int wstride,
                                           const int nw, const int lpv_sh, const int NP, const int la, KP kind,
                                           const T* __restrict__ dtdp, float* __restrict__ part, const unsigned m) {
    constexpr int np = NPhase<T>::value;
    constexpr int wsh = W == 2 ? 1 : W == 4 ? 2 : W == 8 ? 3 : 4;  // log2 W
    const int NPh = NP / 2;
    typedef typename CoefOf<T, KF>::type CT;
    constexpr int NC = np * CoefOf<T, KF>::per_param;
    const CT* src = reinterpret_cast<const CT*>(dtdp) + (size_t)la * NPh * NC;
    it.out = part + ((size_t)la * NPh) * np + m * np;
    if constexpr (KF) {
        // Slot-major entries; inactive slots have zero coefficients.
        it.state = 2;
#pragma unroll
        for (int wv = 0; wv < NW; ++wv)
            if (wv < nw) it.a[wv] = s0[wv * wstride + m];
#pragma unroll
        for (int e = 0; e < NC; ++e) it.dd[e] = (src + (size_t)e * NPh)[m];
        return;
    }
    const int kd = kind[la];
    int i, j;
    if (kd <= 1) {
        i = 2 * (int)m + kd;
        j = i + 1;
    } else if (kd != KIND_IDENTITY) {
        const int st = kd - 2;
        i = (((int)m >> st) << (st + 1)) + ((int)m & ((1 << st) - 1));
        j = i + (1 << st);
    } else {
        i = -1;
        j = 0;
    }
    if (i < 0 || j >= NP) {
        it.state = 1;
        return;
    }
    it.state = 2;
    // Stage index of waveguide i: (i % W) * LPV + i / W, LPV = 2^lpv_sh.
    const int oi = ((i & (W - 1)) << lpv_sh) + (i >> wsh), oj = ((j & (W - 1)) << lpv_sh) + (j >> wsh);
#pragma unroll
    for (int wv = 0; wv < NW; ++wv)
        if (wv < nw) {
            it.a[wv] = s0[wv * wstride + oi];
            it.b[wv] = s0[wv * wstride + oj];
        }
    // One uniform base per table row, one 32-bit lane offset for all of them.
#pragma unroll
    for (int e = 0; e < NC; ++e) it.dd[e] = (src + (size_t)e * NPh)[m];
}

template <class T, int NW, bool KF>
__device__ __forceinline__ void conv_finish(const ConvItem<T, NW, KF>& it, const int nw, const bool first_tile) {
    typedef typename RowOf<T>::type R;
    constexpr int np = NPhase<T>::value;
    if (it.state == 0) return;
    if (it.state == 1) {
        if (first_tile)
#pragma unroll
            for (int k = 0; k < np; ++k) it.out[k] = 0.f;
        return;
    }
    float val[np];
    if constexpr (KF) {
        R e = it.a[0];
#pragma unroll
        for (int wv = 1; wv < NW; ++wv)
            if (wv < nw) e = row_add(e, it.a[wv]);
#pragma unroll
        for (int k = 0; k < np; ++k) val[k] = kdot(e, it.dd[k]);
    } else {
        R si = it.a[0], sj = it.b[0];
#pragma unroll
        for (int wv = 1; wv < NW; ++wv)
            if (wv < nw) {
                si = row_add(si, it.a[wv]);
                sj = row_add(sj, it.b[wv]);
            }
        slot_dp<T, np>(si, sj, it.dd, val);
    }
    put_dp<np>(it.out, val, first_tile);
}

template <class T, int W, int V, bool UNI, int NW, class KP, bool SC = false>
__device__ __forceinline__ void bwd_tile_pipe(const int N, const int l0, const int nl, const int B, const int LPV,
                                              const int NP, const int C, KP kind,
                                              const typename RowOf<T>::type* __restrict__ rows_inv,
                                              const typename RowOf<T>::type* __restrict__ rows_adj,
                                              const T* __restrict__ dtdp, float* __restrict__ my_partial,
                                              const T* u_out, const T* g_out, const T* target, T* u_in, T* g_in,
                                              const int ldu, const long tile, const bool first,
                                              typename RowOf<T>::type* __restrict__ stage,
                                              const Ckpt ck = Ckpt{nullptr, 0, 0},
                                              const float* __restrict__ scales = nullptr) {
    typedef typename RowOf<T>::type R;
    constexpr int D = 2;  // row register sets: rows requested one layer ahead
    const int nw = blockDim.x >> 6;
    const int NPh = NP / 2;
    const int wave_u = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lrows = UNI ? NPh : NP;  // stage entries per wave and layer (K-form: one per slot)
    const int wstride = C * lrows;     // one wave's rows in one buffer
    const int bstride = nw * wstride;  // one buffer
    const int nthr = nw * 64;
    const KP kp = kind + l0;
    const int last = nl - 1;
    const Lanes q = lanes_of<V>(LPV, tile);
    T ua[V][W], ga[V][W], ub[V][W], gb[V][W];
    load_tile(ua, u_out, q, N, B, ldu);
    if (target != nullptr) {
        T y[V][W];
        load_tile(y, target, q, N, B, ldu);
#pragma unroll
        for (int v = 0; v < V; ++v)
#pragma unroll
            for (int k = 0; k < W; ++k) ga[v][k] = fsub2(ua[v][k], y[v][k]);
    } else {
        load_tile(ga, g_out, q, N, B, ldu);
    }
    const R* rip = rows_inv + (size_t)l0 * NP + q.t;
    const R* rap = rows_adj + (size_t)l0 * NP + q.t;
    R ri[D][W], ra[D][W];
    float sc[SC ? D : 1][SC ? W : 1];  // per-waveguide adjoint factors (scaled-unitary crossings)
    const float* scp = SC ? scales + (size_t)l0 * NP + q.t : nullptr;
    int kk[D];
    auto lrow = [&](const int s) { return (size_t)(s < last ? last - s : 0) * NP; };
    if (nl > 0) {
        kk[0] = kp[s_clamp(last, 0)];
        load_rows(ri[0], rip + lrow(0), LPV);
        if constexpr (SC) load_scl(sc[0], scp + lrow(0), LPV);
        if constexpr (!UNI) load_rows(ra[0], rap + lrow(0), LPV);
    }
#define MP_PIPE_LAYER(KD, U, G, UO, GO, RI, RA, LC, SCL)                                \
    {                                                                                   \
        if constexpr (UNI)                                                              \
            bwd_layer<T, W, V, UNI, SC>(KD, U, G, UO, GO, RI, RI, buf, q, LPV, C, LC, SCL); \
        else                                                                            \
            bwd_layer<T, W, V, UNI>(KD, U, G, UO, GO, RI, RA, buf, q, LPV, C, LC);      \
    }
    // Conversion of the previous stage: items idx = gthr + k * nthr < pcl * NPh, layer idx / NPh
    // of that stage (the same for the whole wave), slot idx % NPh.
    int pc0 = 0, pcl = 0;
    R* pbuf = stage;
    // NP is a power of two (a power-of-two lane count times W): divisions by LPV and NP/2 are shifts.
    const int lpv_sh = __builtin_ctz(LPV), nph_sh = __builtin_ctz(NPh);
    // Item k of this thread (none when k is past the previous stage's items).
    ConvItem<T, NW, UNI> item;
    item.state = 0;
    auto issue = [&](const int k) {
        item.state = 0;
        const int wbase = wave_u * 64 + k * nthr;  // first item of this wave (uniform)
        if (wbase < pcl * NPh) {
            const int lcc = wbase >> nph_sh;
            const int la = l0 + last - (pc0 + lcc);
            const unsigned m = (unsigned)((wbase & (NPh - 1)) + (threadIdx.x & 63));
            conv_issue<T, W, NW, UNI, KP>(item, pbuf + lcc * lrows, wstride, nw, lpv_sh, NP, la, kind, dtdp, my_partial,
                                     m);
        }
    };
    int sidx = 0;
    for (int c0 = 0; c0 < nl; c0 += C, ++sidx) {
        const int cl = (C < nl - c0) ? C : nl - c0;
        R* buf = stage + (sidx & 1) * bstride;
        if (ck.base != nullptr && c0 > 0) {
            const int e = l0 + nl - c0;
            if (e % ck.every == 0)
                load_tile(ua, reinterpret_cast<const T*>(ck.base) + (size_t)(e / ck.every) * ck.stride, q, N, B, ldu);
        }
        const int nitem = (pcl * NPh + nthr - 1) / nthr;  // conversion items per thread (wave-uniform)
        int lc = 0;
        for (; lc + D <= cl; lc += D) {
            const int s = c0 + lc;
            if (lc < nitem) issue(lc);
            kk[1] = kp[s_clamp(last, s + 1)];
            load_rows(ri[1], rip + lrow(s + 1), LPV);
            if constexpr (SC) load_scl(sc[1 % (SC ? D : 1)], scp + lrow(s + 1), LPV);
            if constexpr (!UNI) load_rows(ra[1], rap + lrow(s + 1), LPV);
            MP_PIPE_LAYER(kk[0], ua, ga, ub, gb, ri[0], ra[0], lc, sc[0])
            if (lc < nitem) conv_finish(item, nw, first);
            if (lc + 1 < nitem) issue(lc + 1);
            kk[0] = kp[s_clamp(last, s + 2)];
            load_rows(ri[0], rip + lrow(s + 2), LPV);
            if constexpr (SC) load_scl(sc[0], scp + lrow(s + 2), LPV);
            if constexpr (!UNI) load_rows(ra[0], rap + lrow(s + 2), LPV);
            MP_PIPE_LAYER(kk[1], ub, gb, ua, ga, ri[1], ra[1], lc + 1, sc[1 % (SC ? D : 1)])
            if (lc + 1 < nitem) conv_finish(item, nw, first);
        }
        if (lc < cl) {  // one layer left (last stage of an odd depth)
            if (lc < nitem) issue(lc);
            MP_PIPE_LAYER(kk[0], ua, ga, ub, gb, ri[0], ra[0], lc, sc[0])
            if (lc < nitem) conv_finish(item, nw, first);
            copy_state(ub, ua);
            copy_state(gb, ga);
            ++lc;
        }
        for (int k = lc; k < nitem; ++k) {  // items beyond the stage's layers
            issue(k);
            conv_finish(item, nw, first);
        }
        __syncthreads();
        pc0 = c0;
        pcl = cl;
        pbuf = buf;
    }
    // The last stage: converted after its barrier, without overlap.
    {
        const int nitem = (pcl * NPh + nthr - 1) / nthr;
        for (int k = 0; k < nitem; ++k) {
            issue(k);
            conv_finish(item, nw, first);
        }
    }
    __syncthreads();  // the next tile of this workgroup reuses the buffers
#undef MP_PIPE_LAYER
    if (u_in != nullptr) store_tile(ua, u_in, q, N, B, ldu);
    if (g_in != nullptr) store_tile(ga, g_in, q, N, B, ldu);
}

// ---- neighbour-mesh adjoint with a compile-time stage ------------------------------------
// The pipelined unitary adjoint for NN geometries (layer kinds 0 and 1 only) whose vectors
// fill one wave (LPV = 64, so NP = 64 W), compiled with the workgroup's wave count NW and the
// stage depth C as constants.  Same arithmetic and the same fixed summation orders as
// bwd_tile_pipe (its results are bitwise equal), with the run-time structure of the general
// tile removed from the layer loop: no lane-group folds, no FFT stage-write variants, stage
// offsets and conversion items known at compile time (the C layers of a stage are unrolled),
// the C layer kinds of a stage read once at its start, and the rows (and scaled-unitary
// factors) read with raw buffer loads (scalar layer offset, fixed lane offset).
//
// Stage order of the NN-specialised adjoint of width W.  Slot order puts lane t's W/2 entries
// side by side (lane stride W/2 entries of 16 B: bank conflicts on the stage writes grow with
// W); stage order puts entry p of lane t at p*64 + t, so each write instruction of a wave is
// 1 KiB of consecutive LDS.  The conversion, the K-form table (written by the prologue in the
// same order) and the dp partials are then indexed by stage position, and the dp reduction
// maps positions back to slots.
// Measured against slot order (MI355X, fwd+grad, Clements MZI): N=256 L=256 B=4096 0.2890 ms
// against 0.2916 ms; N=512 L=512 B=4096 1.483 against 1.515 ms; at W = 16 slot order made the
// one-layer stage spill and conflict (W = 2 has one slot per lane: both orders agree).
__host__ __device__ constexpr bool nn_stage_order(int W) { return W >= 4; }

// Whether the NN-specialised adjoint stages its rows through LDS: eight-wave workgroups, W >= 8
// (scaled-unitary adjoints stage the rows and load their per-waveguide factors per wave, one
// layer ahead).  Measured against per-wave row loads (fwd+grad, Clements MZI, L = N): N=1024
// B=2048 3.91 -> 3.54 ms, N=512 B=4096 1.487 -> 1.404 ms; the lossy N=512 L=512 B=8192 Adam
// step with checkpoints every 16 layers 4.56 -> 4.28 ms; at W = 4 (N=256 B=4096, two vectors
// per lane) 0.2891 -> 0.2942 ms, so it stays off there.
__host__ __device__ constexpr bool nn_rows_staged(int W, int NW, bool) { return W >= 8 && NW == 8; }

// K-form stage entries of one NN layer (lane t owns slots t*W/2 .. t*W/2 + W/2 - 1; an odd
// layer pairs the lane's last waveguide with the next lane's first); entry p at wl[p * ES].
template <class T, int W, int ES>
__device__ __forceinline__ void kstage_write_nn(const bool odd, const T (&a)[W], const T (&b)[W],
                                                typename RowOf<T>::type* wl) {
    constexpr int H = W / 2;
    if (odd) {
        const T na = a_next(a[0]), nb = lane_next(b[0]);
#pragma unroll
        for (int p = 0; p + 1 < H; ++p)
            wl[p * ES] = kquad(a[2 * p + 1], a[2 * p + 2], b[2 * p + 1], b[2 * p + 2]);
        wl[(H - 1) * ES] = kquad(a[W - 1], na, b[W - 1], nb);
    } else {
#pragma unroll
        for (int p = 0; p < H; ++p) wl[p * ES] = kquad(a[2 * p], a[2 * p + 1], b[2 * p], b[2 * p + 1]);
    }
}

// One NN layer of the unitary adjoint (bwd_layer without folds or kind dispatch).
template <class T, int W, int V, bool SC>
__device__ __forceinline__ void bwd_layer_nn(const bool odd, const T (&u)[V][W], const T (&g)[V][W], T (&uo)[V][W],
                                             T (&go)[V][W], const typename RowOf<T>::type (&ri)[W],
                                             typename RowOf<T>::type* wl, const int lane, const float* scl) {
    T accA[W], accB[W];
#pragma unroll
    for (int k = 0; k < W; ++k) {
        accA[k] = fzero(T());
        accB[k] = fzero(T());
    }
    BwdStep<T, W, V, true, SC> st{u, g, uo, go, ri, ri, accA, accB, lane, scl};
    if (odd)
        st.template run<1, 1>(0);
    else
        st.template run<0, 1>(0);
    kstage_write_nn<T, W, nn_stage_order(W) ? 64 : 1>(odd, accA, accB, wl);
}

template <class T, int W, int V, int NW, int C, bool SC, bool RSX = nn_rows_staged(W, NW, SC)>
__device__ __forceinline__ void bwd_tile_nn(const int N, const int l0, const int nl, const int B, ConstInt* kind,
                                            const typename RowOf<T>::type* __restrict__ rows_inv,
                                            const T* __restrict__ dtdp, float* __restrict__ my_partial,
                                            const T* u_out, const T* g_out, const T* target, T* u_in, T* g_in,
                                            const int ldu, const long tile, const bool first,
                                            typename RowOf<T>::type* __restrict__ stage, const Ckpt ck,
                                            const float* __restrict__ scales) {
    typedef typename RowOf<T>::type R;
    typedef typename KOf<T>::type KT;
    constexpr int LPV = 64, NP = 64 * W, NPh = NP / 2;
    constexpr int np = NPhase<T>::value;
    constexpr int NTHR = 64 * NW;
    constexpr int ITEMS = C * NPh;                     // conversion items of a whole stage
    constexpr int NI = (ITEMS + NTHR - 1) / NTHR;      // items per thread and stage
    constexpr int WSTRIDE = C * NPh;                   // one wave's entries in one buffer
    constexpr int BSTRIDE = NW * WSTRIDE;              // one buffer
    // The field and row register sets alternate per layer (set of local layer s: s & 1); a
    // stage of odd depth starts on either set, so such stages run in pairs.
    static_assert(C % 2 == 0 || C == 1, "stage depth: even, or one layer");
    static_assert(C > 1 || nn_stage_order(W), "one-layer stages are compiled for the stage order only");
    // Widest tile: the rows are loaded right before their layer (one row set live).
    constexpr bool LEAN = W >= 16;
    static_assert(!(LEAN && NW == 8) || C == 1, "staged rows: one layer per stage");
    const int lane = threadIdx.x & 63;
    const int wave_u = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int last = nl - 1;
    const Lanes q = lanes_of<V>(LPV, tile);
    T ua[V][W], ga[V][W], ub[V][W], gb[V][W];
    load_tile(ua, u_out, q, N, B, ldu);
    if (target != nullptr) {
        T y[V][W];
        load_tile(y, target, q, N, B, ldu);
#pragma unroll
        for (int v = 0; v < V; ++v)
#pragma unroll
            for (int k = 0; k < W; ++k) ga[v][k] = fsub2(ua[v][k], y[v][k]);
    } else {
        load_tile(ga, g_out, q, N, B, ldu);
    }
    // Rows of local layer s (global layer l0 + last - s; past the end: layer l0) by buffer loads.
    const __amdgpu_buffer_rsrc_t rrs = table_rsrc(rows_inv + (size_t)l0 * NP);
    const __amdgpu_buffer_rsrc_t srs = table_rsrc(SC ? scales + (size_t)l0 * NP : nullptr);
    const unsigned roff = (unsigned)lane * (unsigned)sizeof(R), soff = (unsigned)lane * (unsigned)sizeof(float);
    R ri[2][W];
    float sc[2][SC ? W : 1];
    auto rows_at = [&](const int s, R (&r)[W], float (&f)[SC ? W : 1]) {
        const int lr = s < last ? last - s : 0;
#pragma unroll
        for (int k = 0; k < W; ++k) r[k] = buf_load<R>(rrs, roff, (lr * NP + k * LPV) * (int)sizeof(R));
        if constexpr (SC) {
#pragma unroll
            for (int k = 0; k < W; ++k) f[k] = buf_load<float>(srs, soff, (lr * NP + k * LPV) * (int)sizeof(float));
        }
    };
    // Eight-wave workgroups of the wide tiles stage the rows through LDS as the forward does
    // (k_fwd_nn): every wave reading a layer's rows through the address unit (16 KiB per wave
    // and layer at W = 16) bounds the kernel.  Two slots of C layers behind the stage buffers:
    // during stage k each thread requests its share of the rows of stage k + 1 and writes them
    // to slot (k+1) & 1 at the end of the stage (that slot was last read during stage k - 1,
    // before its barrier); the stage barrier publishes them.
    constexpr bool RS = RSX;
    // Row entries each thread copies per layer; with fewer entries than threads (W = 4 in eight
    // waves) the first NP threads copy one each.
    constexpr int CP = (RS && NP >= NTHR) ? NP / NTHR : 1;
    static_assert(!RS || NP % NTHR == 0 || NP < NTHR, "row copy: whole entries per thread");
    const bool copier = NP >= NTHR || (int)threadIdx.x < NP;
    R* const lrow = stage + 2 * BSTRIDE;    // [2][C][NP] (RS only)
    R cp[C][CP];
    auto fetch_rows = [&](R (&c)[CP], const int s) {
        if (!copier) return;
        const int lr = s < last ? last - s : 0;
#pragma unroll
        for (int e = 0; e < CP; ++e)
            c[e] = buf_load<R>(rrs, (unsigned)threadIdx.x * (unsigned)sizeof(R), (lr * NP + e * NTHR) * (int)sizeof(R));
    };
    auto publish_rows = [&](const int k) {  // the fetched rows of stage k into its slot
        if (!copier) return;
        R* dst = lrow + (k & 1) * (C * NP) + threadIdx.x;
#pragma unroll
        for (int lc = 0; lc < C; ++lc)
#pragma unroll
            for (int e = 0; e < CP; ++e) dst[lc * NP + e * NTHR] = cp[lc][e];
    };
    if constexpr (RS) {
        if (nl > 0) {
#pragma unroll
            for (int lc = 0; lc < C; ++lc) fetch_rows(cp[lc], lc);
            publish_rows(0);
            __syncthreads();
        }
    }
    // Conversion item of the previous stage (k-th of this thread): the NW waves' entries of one
    // slot, its K-form coefficients and the partial it lands in.
    // Eight-wave workgroups of the widest tile read an item's stage entries after the layer
    // instead of before it: eight entries held across the layer spill its registers.
    constexpr bool LATE = LEAN;
    struct Item {
        R a[LATE ? 1 : NW];
        KT dd[np];
        float* out;
        const R* s0;
        bool on;
    } it;
    int pc0 = 0;
    R* pbuf = stage;
    auto issue = [&](const int k, const int pitems) {
        const int wbase = wave_u * 64 + k * NTHR;  // first item of this wave (uniform)
        it.on = wbase < pitems;
        if (!it.on) return;
        const int lcc = wbase / NPh;  // stage layer of the wave's items
        const int m = (wbase % NPh) + lane;
        const int la = l0 + last - (pc0 + lcc);
        const R* s0 = pbuf + lcc * NPh + m;
        if constexpr (LATE) {
            it.s0 = s0;
        } else {
#pragma unroll
            for (int wv = 0; wv < NW; ++wv) it.a[wv] = s0[wv * WSTRIDE];
        }
        const KT* src = reinterpret_cast<const KT*>(dtdp) + (size_t)la * np * NPh + m;
#pragma unroll
        for (int e = 0; e < np; ++e) it.dd[e] = src[e * NPh];
        it.out = my_partial + ((size_t)la * NPh + m) * np;
    };
    auto finish = [&]() {
        if (!it.on) return;
        R a[NW];
#pragma unroll
        for (int wv = 0; wv < NW; ++wv) {
            if constexpr (LATE)
                a[wv] = it.s0[wv * WSTRIDE];
            else
                a[wv] = it.a[wv];
        }
        R e = a[0];
#pragma unroll
        for (int wv = 1; wv < NW; ++wv) e = row_add(e, a[wv]);
        float val[np];
#pragma unroll
        for (int k = 0; k < np; ++k) val[k] = kdot(e, it.dd[k]);
        put_dp<np>(it.out, val, first);
    };
    // Per-waveguide factors of a scaled-unitary adjoint with staged rows: one layer ahead,
    // by buffer loads (the one-layer stages of the widest tile load them right before the layer).
    auto scales_at = [&](const int s, float (&f)[SC ? W : 1]) {
        if constexpr (SC) {
            const int lr = s < last ? last - s : 0;
#pragma unroll
            for (int k = 0; k < W; ++k) f[k] = buf_load<float>(srs, soff, (lr * NP + k * LPV) * (int)sizeof(float));
        }
    };
    if (!LEAN && !RS && nl > 0) rows_at(0, ri[0], sc[0]);
    if (RS && SC && !LEAN && nl > 0) scales_at(0, sc[0]);
    bool have_prev = false;
    // One stage of cl layers starting at local layer c0 (GUARD: cl < C, the last stage); PAR is
    // the parity of c0, which picks the register sets of its layers; ODD is the shift parity of
    // a one-layer stage when known at compile time (-1: read per layer).
    auto run_stage = [&](auto guard, auto par, auto odd_c, const int c0, const int cl, R* buf) {
        constexpr bool GUARD = decltype(guard)::value;
        constexpr int PAR = decltype(par)::value;
        constexpr int ODD = decltype(odd_c)::value;
        if (ck.base != nullptr && c0 > 0) {
            const int e = l0 + nl - c0;
            if (e % ck.every == 0) {
                const T* cp = reinterpret_cast<const T*>(ck.base) + (size_t)(e / ck.every) * ck.stride;
                if constexpr (PAR == 0)
                    load_tile(ua, cp, q, N, B, ldu);
                else
                    load_tile(ub, cp, q, N, B, ldu);
            }
        }
        int kd[C];
#pragma unroll
        for (int lc = 0; lc < C; ++lc) kd[lc] = kind[l0 + (c0 + lc < last ? last - (c0 + lc) : 0)];
        R* wl = buf + wave_u * WSTRIDE + (nn_stage_order(W) ? lane : lane * (W / 2));
#pragma unroll
        for (int lc = 0; lc < C; ++lc) {
            if (!GUARD || lc < cl) {
                const int cur = (lc + PAR) & 1;  // register sets of this layer (a constant)
                const int rs = LEAN ? 0 : cur;
                if (lc < NI && have_prev) issue(lc, ITEMS);
                if constexpr (RS) {
                    const R* lr = lrow + (((c0 / C) & 1) * C + lc) * NP + lane;
#pragma unroll
                    for (int k = 0; k < W; ++k) ri[rs][k] = lr[k * LPV];
                    if constexpr (SC) {
                        if constexpr (LEAN)
                            scales_at(c0 + lc, sc[0]);
                        else
                            scales_at(c0 + lc + 1, sc[cur ^ 1]);
                    }
                    fetch_rows(cp[lc], c0 + C + lc);
                } else if constexpr (LEAN) {
                    rows_at(c0 + lc, ri[0], sc[0]);
                } else {
                    rows_at(c0 + lc + 1, ri[cur ^ 1], sc[cur ^ 1]);
                }
                const bool odd = ODD >= 0 ? ODD != 0 : (kd[lc] & 1) != 0;
                if constexpr (C == 1)  // one-layer stages update the state in place
                    bwd_layer_nn<T, W, V, SC>(odd, ua, ga, ua, ga, ri[rs], wl, lane, sc[rs]);
                else if (cur == 0)
                    bwd_layer_nn<T, W, V, SC>(odd, ua, ga, ub, gb, ri[rs], wl + lc * NPh, lane, sc[rs]);
                else
                    bwd_layer_nn<T, W, V, SC>(odd, ub, gb, ua, ga, ri[rs], wl + lc * NPh, lane, sc[rs]);
                if (lc < NI && have_prev) finish();
            }
        }
        if constexpr (RS) publish_rows(c0 / C + 1);
        if (have_prev) {
            // Items of the previous stage that found no layer to hide under.
            const int done = GUARD ? (cl < NI ? cl : NI) : (C < NI ? C : NI);
            for (int k = done; k < NI; ++k) {
                issue(k, ITEMS);
                finish();
            }
        }
        __syncthreads();
        pc0 = c0;
        pbuf = buf;
        have_prev = true;
    };
    typedef std::integral_constant<int, 0> I0;
    typedef std::integral_constant<int, 1> I1;
    typedef std::integral_constant<int, -1> Irt;
    int c0 = 0, sidx = 0;
    if constexpr (C % 2 == 0) {
        for (; c0 + C <= nl; c0 += C, ++sidx)
            run_stage(std::false_type(), I0(), Irt(), c0, C, stage + (sidx & 1) * BSTRIDE);
    } else {
        // One-layer stages, the state updated in place, one body per shift parity.
        for (; c0 < nl; ++c0, ++sidx) {
            R* buf = stage + (sidx & 1) * BSTRIDE;
            if (kind[l0 + last - c0] & 1)
                run_stage(std::false_type(), I0(), I1(), c0, 1, buf);
            else
                run_stage(std::false_type(), I0(), I0(), c0, 1, buf);
        }
    }
    int pitems = ITEMS;
    if constexpr (C > 1) {
        if (c0 < nl) {
            run_stage(std::true_type(), I0(), Irt(), c0, nl - c0, stage + (sidx & 1) * BSTRIDE);
            pitems = (nl - c0) * NPh;
        }
    }
    // An odd number of layers leaves the state in the second register set.
    if (C > 1 && (nl & 1)) {
        copy_state(ub, ua);
        copy_state(gb, ga);
    }
    // The last stage: converted after its barrier, without overlap.
    if (have_prev) {
        const int ni = (pitems + NTHR - 1) / NTHR;
        for (int k = 0; k < ni; ++k) {
            issue(k, pitems);
            finish();
        }
    }
    __syncthreads();  // the next tile of this workgroup reuses the buffers
    if (u_in != nullptr) store_tile(ua, u_in, q, N, B, ldu);
    if (g_in != nullptr) store_tile(ga, g_in, q, N, B, ldu);
}

#define MP_BWD_NN_PARAMS                                                                                  \
    int N, int l0, int nl, int B, const int *__restrict__ kind, const typename RowOf<T>::type *__restrict__ rows_inv, \
        const T *__restrict__ dtdp, float *__restrict__ partial, size_t part_stride, long ntiles, const T *u_out,   \
        const T *g_out, const T *target, T *u_in, T *g_in, int ldu, Ckpt ck, const float *__restrict__ scales
template <class T, int W, int V, int NW, int C, bool SC, bool RSX>
__device__ __forceinline__ void bwd_nn_body(MP_BWD_NN_PARAMS) {
    typedef typename RowOf<T>::type R;
    extern __shared__ unsigned char smem_raw[];
    R* stage = reinterpret_cast<R*>(smem_raw);  // [2][NW][C][NP/2]
    float* my_partial = partial + (size_t)blockIdx.x * part_stride;
    bool first = true;
    for (long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        bwd_tile_nn<T, W, V, NW, C, SC, RSX>(N, l0, nl, B, (ConstInt*)kind, rows_inv, dtdp, my_partial, u_out,
                                             g_out, target, u_in, g_in, ldu, tile, first, stage, ck, scales);
        first = false;
    }
}
template <class T, int W, int V, int NW, int C, bool SC, bool RSX = nn_rows_staged(W, NW, SC)>
__global__ void __launch_bounds__(64 * NW) k_bwd_nn(MP_BWD_NN_PARAMS) {
    bwd_nn_body<T, W, V, NW, C, SC, RSX>(N, l0, nl, B, kind, rows_inv, dtdp, partial, part_stride, ntiles, u_out,
                                         g_out, target, u_in, g_in, ldu, ck, scales);
}
// The widest tile (W = 16), compiled for two waves per SIMD also in four-wave workgroups.
template <class T, int W, int V, int NW, int C, bool SC, bool RSX = nn_rows_staged(W, NW, SC)>
__global__ void __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(2))) k_bwd_nn_w16(MP_BWD_NN_PARAMS) {
    bwd_nn_body<T, W, V, NW, C, SC, RSX>(N, l0, nl, B, kind, rows_inv, dtdp, partial, part_stride, ntiles, u_out,
                                         g_out, target, u_in, g_in, ldu, ck, scales);
}
#undef MP_BWD_NN_PARAMS

// PIPE: 0 = bwd_tile; 4 or 8 = bwd_tile_pipe for workgroups of up to PIPE waves.
// SC: scaled-unitary crossings, adjoint rows = scales (per waveguide) times rows_inv.
template <class T, int W, int V, bool UNI, int PIPE, bool SC = false>
__global__ void __launch_bounds__(64 * (PIPE ? PIPE : bwd_max_waves(W, V, UNI))) k_bwd(
    int N, int l0, int nl, int B, int LPV, int NP, int C, int np_unused, const int* __restrict__ kind,
    const typename RowOf<T>::type* __restrict__ rows_inv, const typename RowOf<T>::type* __restrict__ rows_adj,
    const T* __restrict__ dtdp, float* __restrict__ partial, size_t part_stride, long ntiles, const T* u_out,
    const T* g_out, const T* target, T* u_in, T* g_in, int ldu, Ckpt ck, const float* __restrict__ scales) {
    typedef typename RowOf<T>::type R;
    extern __shared__ unsigned char smem_raw[];
    // [nw][C][W][LPV] (K-form: [nw][C][NP/2]), twice with PIPE
    R* stage = reinterpret_cast<R*>(smem_raw);
    (void)np_unused;
    float* my_partial = partial + (size_t)blockIdx.x * part_stride;
    bool first = true;
    for (long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        if constexpr (PIPE > 0)
            bwd_tile_pipe<T, W, V, UNI, PIPE, ConstInt*, SC>(N, l0, nl, B, LPV, NP, C, (ConstInt*)kind, rows_inv,
                                                             rows_adj, dtdp, my_partial, u_out, g_out, target, u_in,
                                                             g_in, ldu, tile, first, stage, ck, scales);
        else
            bwd_tile<T, W, V, UNI, ConstInt*, SC>(N, l0, nl, B, LPV, NP, C, (ConstInt*)kind, rows_inv, rows_adj,
                                                  dtdp, my_partial, u_out, g_out, target, u_in, g_in, ldu, tile, first,
                                                  stage, ck, scales);
        first = false;
    }
}

// The same kernel for the widest tile (W = 16), compiled for two waves per SIMD: with the lean
// unitary form its registers fit 256, which the default target (one wave) does not enforce.
template <class T, int W, int V, bool UNI, int PIPE, bool SC = false>
__global__ void __launch_bounds__(64 * (PIPE ? PIPE : bwd_max_waves(W, V, UNI)))
__attribute__((amdgpu_waves_per_eu(2))) k_bwd_w16(
    int N, int l0, int nl, int B, int LPV, int NP, int C, int np_unused, const int* __restrict__ kind,
    const typename RowOf<T>::type* __restrict__ rows_inv, const typename RowOf<T>::type* __restrict__ rows_adj,
    const T* __restrict__ dtdp, float* __restrict__ partial, size_t part_stride, long ntiles, const T* u_out,
    const T* g_out, const T* target, T* u_in, T* g_in, int ldu, Ckpt ck, const float* __restrict__ scales) {
    typedef typename RowOf<T>::type R;
    extern __shared__ unsigned char smem_raw[];
    // [nw][C][W][LPV] (K-form: [nw][C][NP/2]), twice with PIPE
    R* stage = reinterpret_cast<R*>(smem_raw);
    (void)np_unused;
    float* my_partial = partial + (size_t)blockIdx.x * part_stride;
    bool first = true;
    for (long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        if constexpr (PIPE > 0)
            bwd_tile_pipe<T, W, V, UNI, PIPE, ConstInt*, SC>(N, l0, nl, B, LPV, NP, C, (ConstInt*)kind, rows_inv,
                                                             rows_adj, dtdp, my_partial, u_out, g_out, target, u_in,
                                                             g_in, ldu, tile, first, stage, ck, scales);
        else
            bwd_tile<T, W, V, UNI, ConstInt*, SC>(N, l0, nl, B, LPV, NP, C, (ConstInt*)kind, rows_inv, rows_adj,
                                                  dtdp, my_partial, u_out, g_out, target, u_in, g_in, ldu, tile, first,
                                                  stage, ck, scales);
        first = false;
    }
}

// A pointer known to be the same in every lane, moved to scalar registers.
template <class P>
__device__ __forceinline__ P* uniform_ptr(P* p) {
    const unsigned long long v = reinterpret_cast<unsigned long long>(p);
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
    return reinterpret_cast<P*>(((unsigned long long)hi << 32) | lo);
}

// =======================================================================================
// Launch plan (host)
// =======================================================================================
struct Plan {
    int op, X, N, L, B, fft;
    int W, V, LPV, NP, waves, C;
    int pipe;         // adjoint: double-buffered LDS stage with the conversion pipelined (bwd_tile_pipe)
    int np, ns, real;
    long ntiles;
    int grid;         // workgroups of the propagation kernel
    int ncu = 256;    // compute units of the device the plan was made for
    size_t rowbytes;  // sizeof(row) = 16 (complex) or 8 (real)
    size_t off_kind, off_a, off_b, off_dtdp, off_partial, total;  // workspace layout
    int uni;          // crossings are unitary (T^-1 = T^H) or, with scaled, scaled unitaries
    int scaled;       // generic crossings asserted to be c U: adjoint rows = |det T| x inverse rows
    const float* scales = nullptr;  // scaled: per-waveguide |det T| table (set by the caller of a launch)
    int nn = 1;       // the NN-specialised kernels may serve this plan (off: MESHPROP_GENERIC_KERNELS)
    int sorder = 0;   // > 0: K-form table and dp partials in the NN adjoint's stage order (W/2 slots per lane)
};

// Dynamic LDS of the NN-specialised adjoint: two stage buffers of NW waves x C layers x NP/2
// K-form entries, and with staged rows two slots of C layers of NP rows.
__host__ __device__ constexpr size_t nn_lds_bytes(int W, int NW, int C, bool rs, size_t rowbytes) {
    return (size_t)2 * NW * C * (W * 32) * rowbytes + (rs ? (size_t)2 * C * 64 * W * rowbytes : 0);
}

// Four-wave workgroups of the widest tile also stage their rows when the grid leaves at most
// one workgroup per CU (small batches): the rows then arrive from LDS instead of L2, which
// shortens each layer's critical path; with more workgroups the 96 KiB of LDS would cost
// occupancy.
inline bool nn_rows_staged_small(const Plan& P) { return P.W == 16 && P.waves == 4 && P.grid <= P.ncu; }

// Stage depth the NN-specialised adjoint is compiled for: two buffers of C layers of the K-form
// stage fill the 64 KiB (four waves) or 128 KiB (eight waves) budget, C = 16 / W (one layer at
// W = 16).
__host__ __device__ constexpr int nn_stage_depth(int W) { return 16 / W; }

// Whether the NN-specialised forward (k_fwd_nn, rows staged through LDS) serves a plan: NN
// geometry, one lane group per wave, four-wave workgroups, the widest tile (W = 16) and at least
// two waves per SIMD (ncu: the device's CUs).  Measured fwdprop, Clements MZI, L = N (general
// forward -> staged): N=1024 B=2048 1.048 -> 0.700 ms, B=4096 (two vectors per lane) 1.979 ->
// 0.897 ms, N=768 B=2048 0.754 -> 0.635 ms; with one wave per SIMD (N=1024 B=1024) 0.523 ->
// 0.607 ms, and at W = 8 (N=512 B=4096) 0.254 -> 0.277 ms, W = 4 (N=256 B=4096) 0.071 -> 0.087
// ms: there a row feeds two vectors, the address unit is not the limit and the per-layer
// barrier costs more than it saves.
// The NN-specialised forward-mode walk (k_fwddiff_nn) serves W >= 8 NN meshes of one vector
// per wave when the batch gives every CU an eight-wave workgroup (its LDS ring holds three
// layers of both row tables).  Measured against the general tile (MI355X, L = N): MZI N=1024
// B=2048 1.654 ms against 2.345, N=768 1.233 against 1.808, orth N=1024 0.829 against 1.83;
// N=512 B=4096 0.785 against 0.978, B=8192 1.40 against 1.81, N=300 0.454 against 0.579, orth
// N=512 0.373 against 0.505.  Not at N=1024 B=256 (32 workgroups: 1.377 against 1.109) nor at
// W = 4 (one vector per wave in four-wave workgroups: N=256 B=4096 0.246 against 0.214, B=16384
// 0.780 against 0.747, orth 0.119 against 0.101).
inline bool nn_fwddiff(const Plan& P) {
    return P.nn && !P.fft && P.LPV == 64 && P.W >= 8 && (long)P.B >= 8L * P.ncu;
}
inline size_t nn_fwddiff_lds(const Plan& P) { return (size_t)3 * 2 * P.NP * P.rowbytes; }

inline bool nn_forward(const Plan& P, const int ncu) {
    return P.nn && !P.fft && P.LPV == 64 && P.W == 16 && P.waves == 4 && P.ntiles * P.waves >= 8L * ncu;
}

// Whether the NN-specialised adjoint (k_bwd_nn) serves a plan: pipelined unitary (or
// scaled-unitary) adjoint of complex fields on an NN geometry, one lane group per wave, four- or
// eight-wave workgroups and the stage depth it is compiled for (W = 16: one vector per lane).
inline bool nn_adjoint(const Plan& P) {
    return P.nn && P.pipe && !P.fft && P.uni && !P.real && P.LPV == 64 && P.W >= 2 && P.W <= 16 &&
           (P.W < 16 || P.V == 1) && (P.waves == 4 || P.waves == 8) && P.C == nn_stage_depth(P.W);
}

// LDS of one wave's stage layer in the adjoint: one entry per waveguide, or one per crossing
// slot for the K-form of unitary crossings.
// A HIP error left pending on the calling thread by earlier, unrelated work is cleared before
// an entry point's first launch, so the launch checks that follow report this call's own.
inline void clear_pending_hip_error() { (void)hipGetLastError(); }

// Raises a kernel's dynamic LDS limit above the default 64 KiB once per device (the attribute
// persists), and reports a failure instead of launching into it.
inline int lds_limit_once(const void* fn, std::atomic<unsigned long long>& done, size_t shmem) {
    if (shmem <= 64 * 1024) return MESHPROP_OK;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return MESHPROP_ERR_HIP;
    const unsigned long long bit = 1ull << dev;
    if (done.load(std::memory_order_acquire) & bit) return MESHPROP_OK;
    if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem) != hipSuccess)
        return MESHPROP_ERR_HIP;
    done.fetch_or(bit, std::memory_order_acq_rel);
    return MESHPROP_OK;
}

inline size_t stage_layer_bytes(const Plan& P) { return (size_t)(P.uni ? P.NP / 2 : P.NP) * P.rowbytes; }

// Tiles compiled into the library.  fwdprop keeps u[V][W]; fwddiff and backdiff keep two
// field tiles plus row tables, so their W*V is capped lower.
__host__ __device__ constexpr bool fwd_tile_ok(int W, int V) { return W * V <= 32 && V <= 8; }
__host__ __device__ constexpr bool diff_tile_ok(int W, int V) { return W * V <= 16 && V <= 4; }

int launch_fwd_any(const Plan& P, const int* kind, const void* rows, int l0, int nl, const void* u_in, void* u_out,
                   int ldu, hipStream_t st, Ckpt ck = Ckpt{nullptr, 0, 0});
int launch_fwddiff_any(const Plan& P, const int* kind, const void* rows, const void* drows, const void* u_in,
                       const void* du_in, void* u_out, void* du_out, int ldu, hipStream_t st);
int launch_bwd_any(const Plan& P, const int* kind, const void* ri, const void* ra, const void* dtdp, float* partial,
                   size_t part_stride, int l0, int nl, const void* u_out, const void* g_out, const void* target,
                   void* u_in, void* g_in, int ldu, hipStream_t st, Ckpt ck = Ckpt{nullptr, 0, 0});
// dp (+)= the per-workgroup partials (nwg of them, stride E floats) in fixed order; kso > 0:
// partials in the NN adjoint's stage order.  Defined with k_dp_reduce in meshprop.hip.
int launch_dp_reduce(int L, int Nh, int NPh, int np, int nwg, size_t E, const float* partial, float* dp, int ldp,
                     int assign, int kso, hipStream_t st);
// The tiles compiled with an ILP-oriented scheduler (one object per direction).
int launch_fwd_ilp(const Plan& P, const int* kind, const void* rows, int l0, int nl, const void* u_in, void* u_out,
                   int ldu, hipStream_t st, Ckpt ck);
// The scaled-unitary adjoints (their own object: only used for balanced lossy meshes).
int launch_bwd_sc(const Plan& P, const int* kind, const void* ri, const void* ra, const void* dtdp, float* partial,
                  size_t part_stride, int l0, int nl, const void* u_out, const void* g_out, const void* target,
                  void* u_in, void* g_in, int ldu, hipStream_t st, Ckpt ck);
int launch_bwd_ilp(const Plan& P, const int* kind, const void* ri, const void* ra, const void* dtdp, float* partial,
                   size_t part_stride, int l0, int nl, const void* u_out, const void* g_out, const void* target,
                   void* u_in, void* g_in, int ldu, hipStream_t st, Ckpt ck);

}  // namespace mp

