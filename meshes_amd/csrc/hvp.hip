// hvp.hip -- exact Hessian-vector products by forward-over-reverse differentiation (no
// reference counterpart: gpu/mesh.py:236-288 takes central differences of two backdiffs).
//
// Given the forward output and its tangent along the parameter direction q (u_out, du_out,
// from fwddiff), and the loss gradient with its tangent (g, dg; or a target: g = 2 (u - t),
// dg = 2 du), one reverse walk over the layers carries (u, du, g, dg):
//   x  = T^H u,   dx = T^H du + dT_q^H u        (reconstructed input and its tangent)
//   g' = T^H g,   dg' = T^H dg + dT_q^H g      (adjoint and its tangent)
// with dT_q = sum_k q_k dT/dp_k, and keeps the K-form quantities of Q_rc = sum_b conj(g'_r) x_c
// (Im Q00, Im Q11, Q01 - conj(Q10)) and of its tangent dQ = sum_b conj(dg') x + conj(g') dx.
// With K_k = T^H dT/dp_k and its tangent dK_k along q (both anti-Hermitian), every LDS stage
// end converts, per slot,
//   grad_k += Re sum K_k Q,   hvp_k += Re sum (K_k dQ + dK_k Q)
// into per-workgroup partials, reduced afterwards in fixed order.  Unitary crossings: MZI and
// SYM (complex fields) and orth (real fields) on NN meshes up to N = 1024 (k_hvp_nn); MZI and
// SYM on FFT meshes up to N = 512 (k_hvp, one vector per lane group).
#include <algorithm>

#include "kernels.h"

namespace mp {

// Tangent of the K-form coefficients along the parameter direction q, for the exact
// Hessian-vector product of unitary crossings: K_k = T^H dT/dp_k is anti-Hermitian for
// every p, hence so is dK_k = (sum_m q_m dT/dp_m)^H dT/dp_k + T^H S_k with the contracted second
// derivatives S_k = sum_m q_m d2T/dp_k dp_m; stored like K (float4 (a, d, c) per (layer,
// parameter, slot), or the float c of a real crossing; zero for inactive slots).  kso > 0:
// the NN walk's stage order (stage_pos), as its K table.
template <class KT>
__global__ void k_dkform(int X, int N, int L, int NP, int fft, const int* __restrict__ g0, const int* __restrict__ g1,
                         const float* __restrict__ p, const float* __restrict__ q, int ldp, const float* __restrict__ s,
                         int lds, KT* __restrict__ dk, int kso) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    const int l = blockIdx.y;
    const int NPh = NP / 2;
    if (l >= L || m >= NPh) return;
    const int np = n_phase_of(X), ns = n_split_of(X);
    const int km = stage_pos(m, kso);
    KT* out = dk + (size_t)l * np * NPh + km;
    int i;
    if (fft) {
        const int st = g0[l];
        bool ok = st >= 0 && st <= 20;
        const int S = ok ? (1 << st) : 1;
        i = 2 * S * (m / S) + (m % S);
        ok = ok && (i + S < N);
        if (!ok) i = -1;
    } else {
        i = 2 * m + (g1[l] & 1);
    }
    int mm = -1;
    bool upper = false;
    const bool act = (i >= 0) && wg_role(fft, g0, g1, l, i, N, mm, upper) && upper && mm == m;
    if (!act) {
        for (int k = 0; k < np; ++k) out[(size_t)k * NPh] = KT{};
        return;
    }
    const float* pp = p + (size_t)l * ldp + (size_t)m * np;
    const float* qp = q + (size_t)l * ldp + (size_t)m * np;
    const float* sp = (s && ns) ? s + (size_t)l * lds + (size_t)m * ns : nullptr;
    M2 Tm, dT[2], S2[2];
    crossing_eval(X, pp, sp, &Tm, dT);
    const double qq[2] = {qp[0], np > 1 ? (double)qp[1] : 0.0};
    crossing_d2q(X, Tm, dT, qq, S2);
    M2 Dq = mscale(dT[0], qq[0]);
    if (np > 1) Dq = madd(Dq, mscale(dT[1], qq[1]));
    for (int k = 0; k < np; ++k) {
        const M2 dK = madd(mmul(mherm(Dq), dT[k]), mmul(mherm(Tm), S2[k]));
        const double cr = 0.5 * (dK.a01.r - dK.a10.r), ci = 0.5 * (dK.a01.i + dK.a10.i);
        if constexpr (sizeof(KT) == sizeof(float4))
            out[(size_t)k * NPh] = KT{(float)dK.a00.i, (float)dK.a11.i, (float)cr, (float)ci};
        else
            out[(size_t)k * NPh] = KT((float)cr);
    }
}


template <int W>
struct HvpStep {
    const float2 (&u)[W];
    const float2 (&du)[W];
    const float2 (&g)[W];
    const float2 (&dg)[W];
    float2 (&uo)[W];
    float2 (&duo)[W];
    float2 (&go)[W];
    float2 (&dgo)[W];
    const float4 (&r)[W];
    const float4 (&d)[W];
    float2 (&A)[W];
    float2 (&Bq)[W];
    float2 (&dA)[W];
    float2 (&dB)[W];
    int lane;

    __device__ __forceinline__ void identity() {
#pragma unroll
        for (int k = 0; k < W; ++k) {
            uo[k] = u[k];
            duo[k] = du[k];
            go[k] = g[k];
            dgo[k] = dg[k];
        }
    }
    // Partner of waveguide k for layer class KC (in-lane stride S; odd NN shift; lane ^ xm).
    template <int KC, int S>
    __device__ __forceinline__ void partners(const float2 (&a)[W], float2 (&p)[W], const int xm) {
        if constexpr (KC == 0) {
#pragma unroll
            for (int k = 0; k < W; ++k) p[k] = a[k ^ S];
        } else if constexpr (KC == 1) {
#pragma unroll
            for (int k = 1; k + 1 < W; ++k) p[k] = a[(k & 1) ? k + 1 : k - 1];
            p[W - 1] = lane_next(a[0]);
            p[0] = lane_prev(a[W - 1]);
        } else {
            const int src = lane ^ xm;
#pragma unroll
            for (int k = 0; k < W; ++k) p[k] = lane_read(a[k], src);
        }
    }
    template <int KC, int S>
    __device__ __forceinline__ void run(const int xm) {
        float2 pu[W], pdu[W], pg[W], pdg[W];
        partners<KC, S>(u, pu, xm);
        partners<KC, S>(du, pdu, xm);
        partners<KC, S>(g, pg, xm);
        partners<KC, S>(dg, pdg, xm);
#pragma unroll
        for (int k = 0; k < W; ++k) {
            uo[k] = rapply(r[k], u[k], pu[k]);
            duo[k] = rapply_d(r[k], d[k], u[k], pu[k], du[k], pdu[k]);
            go[k] = rapply(r[k], g[k], pg[k]);
            dgo[k] = rapply_d(r[k], d[k], g[k], pg[k], dg[k], pdg[k]);
        }
        float2 px[W], pdx[W];
        partners<KC, S>(uo, px, xm);
        partners<KC, S>(duo, pdx, xm);
        // K-form rows: A = Im(conj(g') x) (.x), B = conj(g') x_partner, and their tangents.
#pragma unroll
        for (int k = 0; k < W; ++k) {
            A[k] = cjim(go[k], uo[k], A[k]);
            Bq[k] = cjfma(go[k], px[k], Bq[k]);
            dA[k] = cjim(go[k], duo[k], cjim(dgo[k], uo[k], dA[k]));
            dB[k] = cjfma(go[k], pdx[k], cjfma(dgo[k], px[k], dB[k]));
        }
    }
};

constexpr int kHvpWaves = 4;
static_assert(kHvpWaves <= kStageWaves, "stage conversion sums at most kStageWaves waves");
constexpr size_t kHvpStage = 64 * 1024;

// Compiled for at least two waves per SIMD: the W = 8 tile otherwise exceeds 256 registers
// by a few and runs one wave per SIMD (the smaller tiles already fit more).
template <int W>
__global__ void __launch_bounds__(64 * kHvpWaves) __attribute__((amdgpu_waves_per_eu(2))) k_hvp(
    int N, int L, int B, int LPV, int NP, int C, const int* __restrict__ kind, const float4* __restrict__ rows_r,
    const float4* __restrict__ rows_d, const float4* __restrict__ ktab, const float4* __restrict__ dktab,
    float* __restrict__ part_g, float* __restrict__ part_h, size_t part_stride, long ntiles, const float2* u_out,
    const float2* du_out, const float2* g_out, const float2* dg_out, const float2* target, int ldu, float2* g_in,
    float2* dg_in) {
    extern __shared__ unsigned char smem_raw[];
    float4* stage = reinterpret_cast<float4*>(smem_raw);  // [nw][C][2][NP/2]: Q entries, then dQ
    const int nw = blockDim.x >> 6;
    const int NPh = NP / 2;
    const int lane = threadIdx.x & 63;
    const int wave_u = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wstride = C * 2 * NPh;
    float* pg = part_g + (size_t)blockIdx.x * part_stride;
    float* ph = part_h + (size_t)blockIdx.x * part_stride;
    bool first = true;
    for (long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const Lanes q = lanes_of<1>(LPV, tile);
        float2 u[1][W], du[1][W], g[1][W], dg[1][W];
        load_tile(u, u_out, q, N, B, ldu);
        load_tile(du, du_out, q, N, B, ldu);
        if (target != nullptr) {
            float2 t[1][W];
            load_tile(t, target, q, N, B, ldu);
#pragma unroll
            for (int k = 0; k < W; ++k) {
                g[0][k] = fsub2(u[0][k], t[0][k]);
                dg[0][k] = make_float2(2.f * du[0][k].x, 2.f * du[0][k].y);
            }
        } else {
            load_tile(g, g_out, q, N, B, ldu);
            load_tile(dg, dg_out, q, N, B, ldu);
        }
        // Layers are walked as s = 0, 1, ... (layer L-1-s) in pairs with two row buffers, so the
        // rows of layer s+1 are in flight while layer s is applied; every C layers the stage is
        // converted.
        // W = 8 keeps one row buffer (loaded right before its layer): two do not fit the
        // registers of the wide tile.
        constexpr bool kPrefetch = W <= 4;
        float4 rA[W], dA_[W], rB[kPrefetch ? W : 1], dB_[kPrefetch ? W : 1];
        auto rows_of = [&](const int s, float4 (&r)[W], float4 (&d)[W]) {
            if (s < L) {
                load_rows(r, lane_ptr(rows_r + (size_t)(L - 1 - s) * NP, q.t), LPV);
                load_rows(d, lane_ptr(rows_d + (size_t)(L - 1 - s) * NP, q.t), LPV);
            }
        };
        auto step = [&](const int s, const float4 (&r)[W], const float4 (&d)[W]) {
            const int kd = kind[L - 1 - s];
            const int lc = s % C;
            float2 uo[W], duo[W], go[W], dgo[W], A[W], Bq[W], dA[W], dB[W];
#pragma unroll
            for (int k = 0; k < W; ++k) A[k] = Bq[k] = dA[k] = dB[k] = make_float2(0.f, 0.f);
            HvpStep<W> st{u[0], du[0], g[0], dg[0], uo, duo, go, dgo, r, d, A, Bq, dA, dB, lane};
            dispatch_kind<W>(kd, st);
            // Fold lane groups (they hold the same waveguides of different vectors).
            for (int off = LPV; off < 64; off <<= 1) {
#pragma unroll
                for (int k = 0; k < W; ++k) {
                    A[k] = fold_im(A[k], lane ^ off);
                    Bq[k] = fadd(Bq[k], lane_read(Bq[k], lane ^ off));
                    dA[k] = fold_im(dA[k], lane ^ off);
                    dB[k] = fadd(dB[k], lane_read(dB[k], lane ^ off));
                }
            }
            // One entry per crossing slot for Q and for dQ (slot-major, every slot written).
            float4* sq = stage + ((size_t)__builtin_amdgcn_readfirstlane(q.wave) * C + lc) * 2 * NPh;
            kstage_write<float2, W>(kd, A, Bq, sq, q);
            kstage_write<float2, W>(kd, dA, dB, sq + NPh, q);
#pragma unroll
            for (int k = 0; k < W; ++k) {
                u[0][k] = uo[k];
                du[0][k] = duo[k];
                g[0][k] = go[k];
                dg[0][k] = dgo[k];
            }
        };
        // Stage end after layer s: convert its cl layers (gradient and HVP of every slot).
        auto convert = [&](const int s) {
            const int c0 = s - s % C, cl = s % C + 1;
            __syncthreads();
            for (int lcc = wave_u; lcc < cl; lcc += nw) {
                const int la = L - 1 - (c0 + lcc);
                const float4* s0 = stage + (size_t)lcc * 2 * NPh;
                const float4* s1 = s0 + NPh;
                const float4* k1 = ktab + (size_t)la * 2 * NPh;
                const float4* k2 = dktab + (size_t)la * 2 * NPh;
                float* og = pg + (size_t)la * NPh * 2;
                float* oh = ph + (size_t)la * NPh * 2;
                for (int m = lane; m < NPh; m += 64) {
                    // Inactive slots have zero coefficients.
                    const float4 K0 = k1[m], K1 = k1[NPh + m], D0 = k2[m], D1 = k2[NPh + m];
                    float4 a_[kStageWaves], b_[kStageWaves];
#pragma unroll
                    for (int wv = 0; wv < kStageWaves; ++wv)
                        if (wv < nw) {
                            a_[wv] = s0[wv * wstride + m];
                            b_[wv] = s1[wv * wstride + m];
                        }
                    float4 Q = a_[0], dQ = b_[0];
#pragma unroll
                    for (int wv = 1; wv < kStageWaves; ++wv)
                        if (wv < nw) {
                            Q = row_add(Q, a_[wv]);
                            dQ = row_add(dQ, b_[wv]);
                        }
                    const float gv[2] = {kdot(Q, K0), kdot(Q, K1)};
                    const float hv[2] = {kdot(dQ, K0) + kdot(Q, D0), kdot(dQ, K1) + kdot(Q, D1)};
                    put_dp<2>(og + m * 2, gv, first);
                    put_dp<2>(oh + m * 2, hv, first);
                }
            }
            __syncthreads();
        };
        if constexpr (kPrefetch) {
            rows_of(0, rA, dA_);
            for (int s = 0; s < L; s += 2) {
                rows_of(s + 1, rB, dB_);
                step(s, rA, dA_);
                if (s % C == C - 1 || s == L - 1) convert(s);
                if (s + 1 < L) {
                    rows_of(s + 2, rA, dA_);
                    step(s + 1, rB, dB_);
                    if ((s + 1) % C == C - 1 || s + 1 == L - 1) convert(s + 1);
                }
            }
        } else {
            for (int s = 0; s < L; ++s) {
                rows_of(s, rA, dA_);
                step(s, rA, dA_);
                if (s % C == C - 1 || s == L - 1) convert(s);
            }
        }
        if (g_in != nullptr) {  // the walk's final gradient and its tangent: those of the input
            store_tile(g, g_in, q, N, B, ldu);
            store_tile(dg, dg_in, q, N, B, ldu);
        }
        first = false;
    }
}

// ---- NN meshes --------------------------------------------------------------------------
// One nearest-neighbour layer of the HVP walk for one vector per wave (lane t holds waveguides
// t*W .. t*W + W - 1), the state updated in place; ODD: the layer pairs (2p+1, 2p+2) in the
// lane and the lane's last waveguide with the next lane's first.  The Q and dQ entries of the
// lane's slot p go to wq[p * 64] and wdq[p * 64] (stage order).
// Row sources of one layer: registers loaded beforehand (RowRegs), or buffer loads issued where
// a crossing needs them (RowLoads; the compiler schedules them against the register budget).
template <class R, int W>
struct RowRegs {
    const R (&ri)[W];
    const R (&rd)[W];
    __device__ __forceinline__ R r(const int k) const { return ri[k]; }
    __device__ __forceinline__ R d(const int k) const { return rd[k]; }
};
template <class R>
struct RowLoads {
    __amdgpu_buffer_rsrc_t rr, dr;
    unsigned off;
    int base;  // byte offset of the layer's k = 0 entries
    int kst;   // bytes between the entries of consecutive k
    __device__ __forceinline__ R r(const int k) const { return buf_load<R>(rr, off, base + k * kst); }
    __device__ __forceinline__ R d(const int k) const { return buf_load<R>(dr, off, base + k * kst); }
};

// A vector split over two waves (halves of 64*W waveguides): the lanes at the split see the
// other half's boundary waveguide (its layer inputs, exchanged through LDS at the layer barrier)
// in place of the neighbour a lane shift would give, and half 0 computes the first waveguide of
// half 1 itself where the crossing between the halves needs its new values (no second barrier).
template <class T>
struct HvpEdge {
    typedef typename RowOf<T>::type R;
    bool hi, lo;     // this lane is half 0's lane 63 / half 1's lane 0
    bool half0;      // wave-uniform: this wave holds half 0
    T u, du, g, dg;  // the other half's boundary waveguide (layer inputs)
    R r, d;          // (half 0) the rows of waveguide 64*W, the first of half 1
};
__device__ __forceinline__ float2 tsel(const bool c, const float2 a, const float2 b) {
    return make_float2(c ? a.x : b.x, c ? a.y : b.y);
}
__device__ __forceinline__ float tsel(const bool c, const float a, const float b) { return c ? a : b; }

template <class T, int W, bool ODD, bool SPLIT, class RS>
__device__ __forceinline__ void hvp_layer_nn(T (&u)[W], T (&du)[W], T (&g)[W], T (&dg)[W], const RS& rs,
                                             typename RowOf<T>::type* wq, typename RowOf<T>::type* wdq,
                                             const HvpEdge<T>& e) {
    typedef typename RowOf<T>::type R;
    constexpr int H = W / 2, ES = 64;
    const T z = fzero(T());
    // In-lane crossing of waveguides k (upper) and j = k + 1, entry p.
    auto pair = [&](const int k, const int j, const int p) __attribute__((always_inline)) {
        const R rk = rs.r(k), rj = rs.r(j), ek = rs.d(k), ej = rs.d(j);
        const T xk = rapply(rk, u[k], u[j]), xj = rapply(rj, u[j], u[k]);
        const T dxk = rapply_d(rk, ek, u[k], u[j], du[k], du[j]);
        const T dxj = rapply_d(rj, ej, u[j], u[k], du[j], du[k]);
        const T gk = rapply(rk, g[k], g[j]), gj = rapply(rj, g[j], g[k]);
        const T dgk = rapply_d(rk, ek, g[k], g[j], dg[k], dg[j]);
        const T dgj = rapply_d(rj, ej, g[j], g[k], dg[j], dg[k]);
        wq[p * ES] = kquad(cjim(gk, xk, z), cjim(gj, xj, z), cjfma(gk, xj, z), cjfma(gj, xk, z));
        wdq[p * ES] = kquad(cjim(gk, dxk, cjim(dgk, xk, z)), cjim(gj, dxj, cjim(dgj, xj, z)),
                            cjfma(gk, dxj, cjfma(dgk, xj, z)), cjfma(gj, dxk, cjfma(dgj, xk, z)));
        u[k] = xk;
        u[j] = xj;
        du[k] = dxk;
        du[j] = dxj;
        g[k] = gk;
        g[j] = gj;
        dg[k] = dgk;
        dg[j] = dgj;
    };
    if constexpr (!ODD) {
#pragma unroll
        for (int p = 0; p < H; ++p) pair(2 * p, 2 * p + 1, p);
    } else {
        // Lane-boundary waveguides: their new values from the neighbours' layer inputs (lanes
        // without a neighbour read 0; their rows are the identity), read before any update.
        T hu = lane_next(u[0]), hdu = lane_next(du[0]), hg = lane_next(g[0]), hdg = lane_next(dg[0]);
        T lu = lane_prev(u[W - 1]), ldu = lane_prev(du[W - 1]), lg = lane_prev(g[W - 1]), ldg = lane_prev(dg[W - 1]);
        if constexpr (SPLIT) {
            hu = tsel(e.hi, e.u, hu);
            hdu = tsel(e.hi, e.du, hdu);
            hg = tsel(e.hi, e.g, hg);
            hdg = tsel(e.hi, e.dg, hdg);
            lu = tsel(e.lo, e.u, lu);
            ldu = tsel(e.lo, e.du, ldu);
            lg = tsel(e.lo, e.g, lg);
            ldg = tsel(e.lo, e.dg, ldg);
        }
        const T uW = u[W - 1], duW = du[W - 1], gW = g[W - 1], dgW = dg[W - 1];
        const R rh = rs.r(W - 1), eh = rs.d(W - 1), rl = rs.r(0), el = rs.d(0);
        const T xh = rapply(rh, u[W - 1], hu), gh = rapply(rh, g[W - 1], hg);
        const T dxh = rapply_d(rh, eh, u[W - 1], hu, du[W - 1], hdu);
        const T dgh = rapply_d(rh, eh, g[W - 1], hg, dg[W - 1], hdg);
        const T xl = rapply(rl, u[0], lu), gl = rapply(rl, g[0], lg);
        const T dxl = rapply_d(rl, el, u[0], lu, du[0], ldu);
        const T dgl = rapply_d(rl, el, g[0], lg, dg[0], ldg);
#pragma unroll
        for (int p = 0; p + 1 < H; ++p) pair(2 * p + 1, 2 * p + 2, p);
        u[W - 1] = xh;
        du[W - 1] = dxh;
        g[W - 1] = gh;
        dg[W - 1] = dgh;
        u[0] = xl;
        du[0] = dxl;
        g[0] = gl;
        dg[0] = dgl;
        // Crossing (W-1 of this lane, 0 of the next): this lane's rows of Q and dQ, and the
        // next lane's (computed there against this lane's reconstructed x, dx).
        T nx = lane_next(xl), ndx = lane_next(dxl);
        const T px = lane_prev(xh), pdx = lane_prev(dxh);
        const T A0 = cjim(gl, xl, z), B0 = cjfma(gl, px, z);
        const T dA0 = cjim(gl, dxl, cjim(dgl, xl, z)), dB0 = cjfma(gl, pdx, cjfma(dgl, px, z));
        T na = a_next(A0), nb = lane_next(B0), nda = a_next(dA0), ndb = lane_next(dB0);
        if constexpr (SPLIT) {
            if (e.half0) {
                // Half 0's lane 63: the first waveguide of half 1, from its layer inputs (hu ..)
                // and this lane's (uW ..), as half 1 computes it.
                const T x5 = rapply(e.r, hu, uW), g5 = rapply(e.r, hg, gW);
                const T dx5 = rapply_d(e.r, e.d, hu, uW, hdu, duW), dg5 = rapply_d(e.r, e.d, hg, gW, hdg, dgW);
                nx = tsel(e.hi, x5, nx);
                ndx = tsel(e.hi, dx5, ndx);
                na = tsel(e.hi, cjim(g5, x5, z), na);
                nb = tsel(e.hi, cjfma(g5, xh, z), nb);
                nda = tsel(e.hi, cjim(g5, dx5, cjim(dg5, x5, z)), nda);
                ndb = tsel(e.hi, cjfma(g5, dxh, cjfma(dg5, xh, z)), ndb);
            }
        }
        wq[(H - 1) * ES] = kquad(cjim(gh, xh, z), na, cjfma(gh, nx, z), nb);
        wdq[(H - 1) * ES] = kquad(cjim(gh, dxh, cjim(dgh, xh, z)), nda, cjfma(gh, ndx, cjfma(dgh, nx, z)), ndb);
    }
}

// HVP walk of NN meshes: NW waves per workgroup, one vector per wave, the layers walked
// backwards with one-layer stages (buffer s & 1 of [2][NW][Q, dQ][NP/2] entries); while layer
// s is applied, the workgroup converts layer s - 1 (its K and dK coefficients requested before
// the layer, its NW entries summed in wave order after it).  Rows: PF, those of the next
// layer requested one layer ahead (two row register sets); JIT, each crossing's rows loaded
// where it is applied; otherwise a whole layer's rows right before it.  part_g may be null (no
// gradient).  SPL = 2: each vector split over two adjacent waves (halves of 64*W waveguides,
// the row tables laid out for 128 lanes per vector, the stage-order positions of half h after
// those of half 0), the boundary waveguides of the halves exchanged through LDS behind the
// stage buffers ([2 layer parities][NW/2 vectors][2 halves][u, du, g, dg]).
template <class T, int W, int NW, bool PF, bool JIT, int SPL = 1>
__device__ __forceinline__ void hvp_nn_body(const int N, const int L, const int B, const int* __restrict__ kind,
                                            const typename RowOf<T>::type* __restrict__ rows_r,
                                            const typename RowOf<T>::type* __restrict__ rows_d,
                                            const typename KOf<T>::type* __restrict__ ktab,
                                            const typename KOf<T>::type* __restrict__ dktab,
                                            float* __restrict__ part_g, float* __restrict__ part_h,
                                            const size_t part_stride, const long ntiles, const T* u_out,
                                            const T* du_out, const T* g_out, const T* dg_out, const T* target,
                                            const int ldu, T* g_in, T* dg_in) {
    typedef typename RowOf<T>::type R;
    typedef typename KOf<T>::type KT;
    static_assert(SPL == 1 || SPL == 2, "one or two waves per vector");
    static_assert(SPL == 1 || (NW % 2 == 0 && !PF), "split vectors: whole vector pairs, rows loaded per layer");
    constexpr bool SPLIT = SPL > 1;
    constexpr int NPW = 64 * W, NPhW = NPW / 2;  // waveguides and slots of one wave
    constexpr int LPVE = 64 * SPL;               // lanes per vector in the row tables
    constexpr int NP = NPW * SPL, NPh = NP / 2;  // of one vector
    constexpr int VPW = NW / SPL;                // vectors per workgroup
    constexpr int np = NPhase<T>::value;
    constexpr int NTHR = 64 * NW;
    constexpr int NI = (NPh + NTHR - 1) / NTHR;  // conversion items per thread and layer
    constexpr int WSTRIDE = 2 * NPhW;            // one wave's Q then dQ entries
    constexpr int BSTRIDE = NW * WSTRIDE;        // one buffer
    static_assert(NPhW % 64 == 0, "whole waves of conversion items");
    extern __shared__ unsigned char smem_raw[];
    R* stage = reinterpret_cast<R*>(smem_raw);
    T* edge = reinterpret_cast<T*>(stage + 2 * BSTRIDE);  // SPLIT: [2][VPW][2][4]
    const int lane = threadIdx.x & 63;
    const int wave_u = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int half = wave_u % SPL, vec = wave_u / SPL;
    float* pg = part_g ? part_g + (size_t)blockIdx.x * part_stride : nullptr;
    float* ph = part_h + (size_t)blockIdx.x * part_stride;
    const __amdgpu_buffer_rsrc_t rrs = table_rsrc(rows_r), drs = table_rsrc(rows_d);
    const unsigned roff = (unsigned)(half * 64 + lane) * (unsigned)sizeof(R);
    auto rows_at = [&](const int s, R (&r)[W], R (&d)[W]) __attribute__((always_inline)) {
        const int la = s < L ? L - 1 - s : 0;
#pragma unroll
        for (int k = 0; k < W; ++k) {
            r[k] = buf_load<R>(rrs, roff, (la * NP + k * LPVE) * (int)sizeof(R));
            d[k] = buf_load<R>(drs, roff, (la * NP + k * LPVE) * (int)sizeof(R));
        }
    };
    HvpEdge<T> eg;
    eg.hi = SPLIT && half == 0 && lane == 63;
    eg.lo = SPLIT && half == 1 && lane == 0;
    eg.half0 = half == 0;
    bool first = true;
    for (long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        Lanes q;
        q.lane = lane;
        q.wave = wave_u;
        q.grp = 0;
        q.t = half * 64 + lane;
        q.vbase = tile * VPW + vec;
        T u[1][W], du[1][W], g[1][W], dg[1][W];
        load_tile(u, u_out, q, N, B, ldu);
        load_tile(du, du_out, q, N, B, ldu);
        if (target != nullptr) {
            T y[1][W];
            load_tile(y, target, q, N, B, ldu);
#pragma unroll
            for (int k = 0; k < W; ++k) {
                g[0][k] = fsub2(u[0][k], y[0][k]);
                dg[0][k] = fscl(du[0][k], 2.f);
            }
        } else {
            load_tile(g, g_out, q, N, B, ldu);
            load_tile(dg, dg_out, q, N, B, ldu);
        }
        // SPLIT: the boundary waveguides of the halves after layer s (s = -1: the loaded state)
        // go to edge slot s & 1; layer s reads slot (s - 1) & 1 after the barrier between them.
        auto put_edge = [&](const int s) __attribute__((always_inline)) {
            if constexpr (SPLIT) {
                T* ed = edge + (((s & 1) * VPW + vec) * 2 + half) * 4;
                if (eg.hi) {
                    ed[0] = u[0][W - 1];
                    ed[1] = du[0][W - 1];
                    ed[2] = g[0][W - 1];
                    ed[3] = dg[0][W - 1];
                }
                if (eg.lo) {
                    ed[0] = u[0][0];
                    ed[1] = du[0][0];
                    ed[2] = g[0][0];
                    ed[3] = dg[0][0];
                }
            }
        };
        if constexpr (SPLIT) {
            put_edge(-1);
            __syncthreads();
        }
        // Conversion items of the previous layer pla (entries in pbuf): item k of a thread is
        // position k*NTHR + threadIdx.x; whole waves are on or off.
        struct Item {
            KT kk[np], dk[np];
        } it[NI];
        int pla = 0;
        const R* pbuf = stage;
        auto issue = [&]() __attribute__((always_inline)) {
#pragma unroll
            for (int k = 0; k < NI; ++k) {
                const int m = k * NTHR + wave_u * 64 + lane;
                if (k * NTHR + wave_u * 64 >= NPh) continue;
                const KT* kp = ktab + (size_t)pla * np * NPh + m;
                const KT* dp = dktab + (size_t)pla * np * NPh + m;
#pragma unroll
                for (int e = 0; e < np; ++e) {
                    it[k].kk[e] = kp[e * NPh];
                    it[k].dk[e] = dp[e * NPh];
                }
            }
        };
        auto finish = [&]() __attribute__((always_inline)) {
#pragma unroll
            for (int k = 0; k < NI; ++k) {
                const int m = k * NTHR + wave_u * 64 + lane;
                if (k * NTHR + wave_u * 64 >= NPh) continue;
                // Position m belongs to half m / NPhW: the entries of that half's waves, summed
                // in vector order.
                const int hf = m / NPhW;
                const R* s0 = pbuf + hf * WSTRIDE + (m - hf * NPhW);
                R Qs = s0[0], dQs = s0[NPhW];
#pragma unroll
                for (int v = 1; v < VPW; ++v) {
                    Qs = row_add(Qs, s0[v * SPL * WSTRIDE]);
                    dQs = row_add(dQs, s0[v * SPL * WSTRIDE + NPhW]);
                }
                float gv[np], hv[np];
#pragma unroll
                for (int e = 0; e < np; ++e) {
                    gv[e] = kdot(Qs, it[k].kk[e]);
                    hv[e] = kdot(dQs, it[k].kk[e]) + kdot(Qs, it[k].dk[e]);
                }
                const size_t o = ((size_t)pla * NPh + m) * np;
                if (pg) put_dp<np>(pg + o, gv, first);
                put_dp<np>(ph + o, hv, first);
            }
        };
        // W >= 8: the coefficients are requested at the conversion (held across the layer
        // they push the tile past the registers of two waves per SIMD).
        constexpr bool LATE = W >= 8;
        auto layer = [&](const int s, const auto& rsrc) __attribute__((always_inline)) {
            R* buf = stage + (s & 1) * BSTRIDE;
            if (s > 0 && !LATE) issue();
            R* wq = buf + wave_u * WSTRIDE + lane;
            const bool odd = (kind[L - 1 - s] & 1) != 0;
            if constexpr (SPLIT) {
                if (odd) {
                    const T* ed = edge + ((((s - 1) & 1) * VPW + vec) * 2 + (1 - half)) * 4;
                    eg.u = ed[0];
                    eg.du = ed[1];
                    eg.g = ed[2];
                    eg.dg = ed[3];
                    if (eg.half0) {
                        const int off = ((L - 1 - s) * NP + 64) * (int)sizeof(R);  // waveguide NPW: k = 0, lane 64
                        eg.r = buf_load<R>(rrs, 0u, off);
                        eg.d = buf_load<R>(drs, 0u, off);
                    }
                }
            }
            if (odd)
                hvp_layer_nn<T, W, true, SPLIT>(u[0], du[0], g[0], dg[0], rsrc, wq, wq + NPhW, eg);
            else
                hvp_layer_nn<T, W, false, SPLIT>(u[0], du[0], g[0], dg[0], rsrc, wq, wq + NPhW, eg);
            if (s > 0) {
                if (LATE) issue();
                finish();
            }
            put_edge(s);
            __syncthreads();
            pla = L - 1 - s;
            pbuf = buf;
        };
        if constexpr (PF) {
            R rA[W], dA[W], rB[W], dB[W];
            rows_at(0, rA, dA);
            for (int s = 0; s < L; s += 2) {
                rows_at(s + 1, rB, dB);
                layer(s, RowRegs<R, W>{rA, dA});
                if (s + 1 < L) {
                    rows_at(s + 2, rA, dA);
                    layer(s + 1, RowRegs<R, W>{rB, dB});
                }
            }
        } else if constexpr (JIT) {
            for (int s = 0; s < L; ++s)
                layer(s, RowLoads<R>{rrs, drs, roff, (L - 1 - s) * NP * (int)sizeof(R), LPVE * (int)sizeof(R)});
        } else {
            for (int s = 0; s < L; ++s) {
                R r[W], d[W];
                rows_at(s, r, d);
                layer(s, RowRegs<R, W>{r, d});
            }
        }
        // The last layer, converted after its barrier.
        if (L > 0) {
            issue();
            finish();
        }
        if (g_in != nullptr) {  // the walk's final gradient and its tangent: those of the input
            store_tile(g, g_in, q, N, B, ldu);
            store_tile(dg, dg_in, q, N, B, ldu);
        }
        __syncthreads();  // the next tile reuses the buffers
        first = false;
    }
}

// The W <= 8 tiles (also the halves of split vectors) compiled for two waves per SIMD.
template <class T, int W, int NW, bool PF, bool JIT, int SPL = 1>
__global__ void __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(2)))
k_hvp_nn(int N, int L, int B, const int* __restrict__ kind, const typename RowOf<T>::type* __restrict__ rows_r,
         const typename RowOf<T>::type* __restrict__ rows_d, const typename KOf<T>::type* __restrict__ ktab,
         const typename KOf<T>::type* __restrict__ dktab, float* __restrict__ part_g, float* __restrict__ part_h,
         size_t part_stride, long ntiles, const T* u_out, const T* du_out, const T* g_out, const T* dg_out,
         const T* target, int ldu, T* g_in, T* dg_in) {
    hvp_nn_body<T, W, NW, PF, JIT, SPL>(N, L, B, kind, rows_r, rows_d, ktab, dktab, part_g, part_h, part_stride,
                                        ntiles, u_out, du_out, g_out, dg_out, target, ldu, g_in, dg_in);
}

}  // namespace mp

using namespace mp;

namespace {

struct HvpPlan {
    int W, LPV, NP, C, grid, nn, nw, kso, np, spl;
    long ntiles;
    size_t E, rowbytes, kbytes, shmem, off_kind, off_r, off_d, off_t1, off_t2, off_pg, off_ph, total;
};

inline size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

// Waves per workgroup of the NN walk: eight (eight vectors, or four split ones, sum their
// entries per conversion; at W = 8 the 128 KiB stage leaves one workgroup per CU, two waves per
// SIMD).  Measured against four at W = 8 (MI355X, Clements L = N): MZI N=512 B=4096 3.77 ms
// against 4.38, N=300 2.10 against 2.45, orth N=512 1.99 against 2.13; B=256 1.52 against 1.47;
// split vectors at N=1024 B=2048 10.4 against 11.4.
constexpr int hvp_nn_waves(int, int) { return 8; }

int hvp_plan(int X, int N, int L, int B, int fft, HvpPlan& P) {
    if (X < 0 || X > 3) return MESHPROP_ERR_ARG;
    if (N <= 0 || (N & 1) || L < 0 || B < 0) return MESHPROP_ERR_ARG;
    if (X == X_GENERIC) return MESHPROP_ERR_UNSUPPORTED;
    P.nn = !fft;
    P.np = n_phase_of(X);
    const bool real = X == X_ORTH;
    P.rowbytes = real ? sizeof(float2) : sizeof(float4);
    P.kbytes = real ? sizeof(float) : sizeof(float4);
    int dev = 0, ncu = 256;
    if (hipGetDevice(&dev) == hipSuccess)
        if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu <= 0) ncu = 256;
    if (P.nn) {
        if (N > 1024) return MESHPROP_ERR_UNSUPPORTED;
        // N > 512: each vector split over two waves of W = 8 (128 lanes per vector).
        P.W = N <= 128 ? 2 : N <= 256 ? 4 : 8;
        P.spl = N > 512 ? 2 : 1;
        P.LPV = 64 * P.spl;
        P.NP = 64 * P.W * P.spl;
        P.nw = hvp_nn_waves(P.W, P.spl);
        P.kso = P.W / 2;
        P.C = 1;
        const int vpw = P.nw / P.spl;  // vectors per workgroup
        P.ntiles = std::max<long>(1, ((long)B + vpw - 1) / vpw);
        P.grid = (int)std::min<long>(P.ntiles, (P.W >= 8 ? 1L : 2L) * ncu);  // W = 8: one per CU (LDS)
        // two stage buffers of [waves][Q, dQ][slots of one wave], and the split vectors' edges
        P.shmem = (size_t)2 * P.nw * 2 * (64 * P.W / 2) * P.rowbytes +
                  (P.spl > 1 ? (size_t)2 * vpw * 2 * 4 * (P.rowbytes / 2) : 0);
    } else {
        if (real || N > 512) return MESHPROP_ERR_UNSUPPORTED;
        P.W = N <= 128 ? 2 : N <= 256 ? 4 : 8;
        int lpv = 1;
        while (lpv < (N + P.W - 1) / P.W) lpv <<= 1;
        P.LPV = lpv;
        P.NP = lpv * P.W;
        P.nw = kHvpWaves;
        P.kso = 0;
        P.spl = 1;
        const int G = 64 / lpv;
        P.ntiles = std::max<long>(1, ((long)B + kHvpWaves * G - 1) / (kHvpWaves * G));
        P.grid = (int)std::min<long>(P.ntiles, 2L * ncu);
        const size_t per_layer = (size_t)kHvpWaves * 2 * (P.NP / 2) * sizeof(float4);  // Q and dQ per slot
        P.C = (int)std::max<size_t>(1, std::min<size_t>(8, kHvpStage / per_layer));
        if ((size_t)P.C * per_layer > kHvpStage) return MESHPROP_ERR_UNSUPPORTED;
        P.shmem = (size_t)P.C * per_layer;
    }
    const size_t NPh = P.NP / 2;
    P.E = (size_t)L * NPh * P.np;
    size_t off = 0;
    P.off_kind = off;
    off = al256(off + (size_t)L * sizeof(int));
    P.off_r = off;
    off = al256(off + (size_t)L * P.NP * P.rowbytes);
    P.off_d = off;
    off = al256(off + (size_t)L * P.NP * P.rowbytes);
    P.off_t1 = off;
    off = al256(off + (size_t)L * NPh * P.np * P.kbytes);
    P.off_t2 = off;
    off = al256(off + (size_t)L * NPh * P.np * P.kbytes);
    P.off_pg = off;
    off = al256(off + (size_t)P.grid * P.E * sizeof(float));
    P.off_ph = off;
    off = al256(off + (size_t)P.grid * P.E * sizeof(float));
    P.total = off;
    return MESHPROP_OK;
}

template <int W>
void launch_hvp(const HvpPlan& P, int N, int L, int B, unsigned char* base, const void* u_out, const void* du_out,
                const void* g, const void* dg, const void* target, int ldu, void* g_in, void* dg_in, hipStream_t st) {
    hipLaunchKernelGGL((k_hvp<W>), dim3(P.grid), dim3(64 * kHvpWaves), P.shmem, st, N, L, B, P.LPV, P.NP, P.C,
                       (const int*)(base + P.off_kind), (const float4*)(base + P.off_r),
                       (const float4*)(base + P.off_d), (const float4*)(base + P.off_t1),
                       (const float4*)(base + P.off_t2), (float*)(base + P.off_pg), (float*)(base + P.off_ph), P.E,
                       P.ntiles, (const float2*)u_out, (const float2*)du_out, (const float2*)g, (const float2*)dg,
                       (const float2*)target, ldu, (float2*)g_in, (float2*)dg_in);
}

template <class T, int W>
int launch_hvp_nn(const HvpPlan& P, int N, int L, int B, unsigned char* base, bool with_grad, const void* u_out,
                  const void* du_out, const void* g, const void* dg, const void* target, int ldu, void* g_in,
                  void* dg_in, hipStream_t st) {
    typedef typename RowOf<T>::type R;
    typedef typename KOf<T>::type KT;
    constexpr int NW = hvp_nn_waves(W, 1);
    constexpr int NW2 = hvp_nn_waves(W, 2);
    // Row sources, as measured (MI355X, Clements L = N): rows of the next layer prefetched at
    // W <= 4 (MZI N=256 B=4096 HVP 1.04 ms against 1.13 with per-crossing loads); per-crossing
    // loads at W = 8 (no spills at two waves per SIMD: N=512 5.77 -> 4.59 ms).
    constexpr bool PF = W <= 4;
    constexpr bool JIT = W == 8;
    static std::atomic<unsigned long long> lds_set{0};
    static std::atomic<unsigned long long> lds_set2{0};
    if constexpr (W == 8) {
        if (P.spl > 1) {
            if (int rc = lds_limit_once((const void*)k_hvp_nn<T, W, NW2, PF, JIT, 2>, lds_set2, P.shmem)) return rc;
            hipLaunchKernelGGL((k_hvp_nn<T, W, NW2, PF, JIT, 2>), dim3(P.grid), dim3(64 * NW2), P.shmem, st, N, L, B,
                               (const int*)(base + P.off_kind), (const R*)(base + P.off_r),
                               (const R*)(base + P.off_d), (const KT*)(base + P.off_t1),
                               (const KT*)(base + P.off_t2), with_grad ? (float*)(base + P.off_pg) : nullptr,
                               (float*)(base + P.off_ph), P.E, P.ntiles, (const T*)u_out, (const T*)du_out,
                               (const T*)g, (const T*)dg, (const T*)target, ldu, (T*)g_in, (T*)dg_in);
            return MESHPROP_OK;
        }
    }
    if (int rc = lds_limit_once((const void*)k_hvp_nn<T, W, NW, PF, JIT>, lds_set, P.shmem)) return rc;
    hipLaunchKernelGGL((k_hvp_nn<T, W, NW, PF, JIT>), dim3(P.grid), dim3(64 * NW), P.shmem, st, N, L, B,
                       (const int*)(base + P.off_kind), (const R*)(base + P.off_r), (const R*)(base + P.off_d),
                       (const KT*)(base + P.off_t1), (const KT*)(base + P.off_t2),
                       with_grad ? (float*)(base + P.off_pg) : nullptr, (float*)(base + P.off_ph), P.E, P.ntiles,
                       (const T*)u_out, (const T*)du_out, (const T*)g, (const T*)dg, (const T*)target, ldu,
                       (T*)g_in, (T*)dg_in);
    return MESHPROP_OK;
}

template <class T>
int launch_hvp_nn_any(const HvpPlan& P, int N, int L, int B, unsigned char* base, bool with_grad, const void* u_out,
                      const void* du_out, const void* g, const void* dg, const void* target, int ldu,
                      void* g_in, void* dg_in, hipStream_t st) {
    switch (P.W) {
        case 2: return launch_hvp_nn<T, 2>(P, N, L, B, base, with_grad, u_out, du_out, g, dg, target, ldu, g_in, dg_in,
                                                 st);
        case 4: return launch_hvp_nn<T, 4>(P, N, L, B, base, with_grad, u_out, du_out, g, dg, target, ldu, g_in, dg_in,
                                                 st);
        default: return launch_hvp_nn<T, 8>(P, N, L, B, base, with_grad, u_out, du_out, g, dg, target, ldu, g_in, dg_in,
                                                 st);
    }
}

}  // namespace

extern "C" {

size_t meshprop_hvp_workspace_size(int X, int N, int L, int B, int fft) {
    HvpPlan P;
    return hvp_plan(X, N, L, B, fft, P) == MESHPROP_OK ? P.total : 0;
}

int meshprop_hvp_grad_in(int X, int N, int L, int B, int fft, const int* geom0, const int* geom1, const float* p,
                         const float* q, int ldp, const float* s, int lds, const void* u_out, const void* du_out,
                         const void* dJdu_out, const void* d_dJdu_out, const void* target, float* grad, float* hvp,
                         void* dJdu_in, void* d_dJdu_in, int ldu, void* workspace, size_t workspace_bytes,
                         void* stream) {
    HvpPlan P;
    int rc = hvp_plan(X, N, L, B, fft, P);
    if (rc) return rc;
    if (B == 0 || L == 0) return MESHPROP_OK;
    if (ldu < N || !u_out || !du_out || !hvp || !p || !q || !geom0 || (!fft && !geom1)) return MESHPROP_ERR_ARG;
    if (!target && (!dJdu_out || !d_dJdu_out)) return MESHPROP_ERR_ARG;
    if (!dJdu_in != !d_dJdu_in) return MESHPROP_ERR_ARG;
    if (ldp < (N / 2) * P.np) return MESHPROP_ERR_ARG;
    const int ns = n_split_of(X);
    if (s && ns && lds < (N / 2) * ns) return MESHPROP_ERR_ARG;
    if (!workspace || workspace_bytes < P.total || (((uintptr_t)workspace) & 255)) return MESHPROP_ERR_WORKSPACE;
    clear_pending_hip_error();
    hipStream_t st = (hipStream_t)stream;
    unsigned char* base = (unsigned char*)workspace;
    const float* sp = ns ? s : nullptr;
    const int tabs = TAB_INV | TAB_DIFF_ADJ | TAB_KUNI;
    const dim3 g2((P.NP / 2 + 255) / 256, L);
    if (X == X_ORTH) {
        hipLaunchKernelGGL((k_rows<float2>), g2, dim3(256), 0, st, X, N, L, P.NP, P.LPV, fft, geom0, geom1, p, ldp,
                           sp, lds, q, tabs, (int*)(base + P.off_kind), (float2*)(base + P.off_r),
                           (float2*)(base + P.off_d), (float*)(base + P.off_t1), nullptr, P.kso);
        if (hipGetLastError() != hipSuccess) return MESHPROP_ERR_HIP;
        hipLaunchKernelGGL(k_dkform<float>, g2, dim3(256), 0, st, X, N, L, P.NP, fft, geom0, geom1, p, q, ldp, sp,
                           lds, (float*)(base + P.off_t2), P.kso);
    } else {
        hipLaunchKernelGGL((k_rows<float4>), g2, dim3(256), 0, st, X, N, L, P.NP, P.LPV, fft, geom0, geom1, p, ldp,
                           sp, lds, q, tabs, (int*)(base + P.off_kind), (float4*)(base + P.off_r),
                           (float4*)(base + P.off_d), (float2*)(base + P.off_t1), nullptr, P.kso);
        if (hipGetLastError() != hipSuccess) return MESHPROP_ERR_HIP;
        hipLaunchKernelGGL(k_dkform<float4>, g2, dim3(256), 0, st, X, N, L, P.NP, fft, geom0, geom1, p, q, ldp, sp,
                           lds, (float4*)(base + P.off_t2), P.kso);
    }
    if (hipGetLastError() != hipSuccess) return MESHPROP_ERR_HIP;
    if (P.nn) {
        rc = (X == X_ORTH)
                 ? launch_hvp_nn_any<float>(P, N, L, B, base, grad != nullptr, u_out, du_out, dJdu_out, d_dJdu_out,
                                            target, ldu, dJdu_in, d_dJdu_in, st)
                 : launch_hvp_nn_any<float2>(P, N, L, B, base, grad != nullptr, u_out, du_out, dJdu_out, d_dJdu_out,
                                             target, ldu, dJdu_in, d_dJdu_in, st);
        if (rc) return rc;
    } else {
        switch (P.W) {
            case 2: launch_hvp<2>(P, N, L, B, base, u_out, du_out, dJdu_out, d_dJdu_out, target, ldu, dJdu_in, d_dJdu_in,
                               st);
                break;
            case 4: launch_hvp<4>(P, N, L, B, base, u_out, du_out, dJdu_out, d_dJdu_out, target, ldu, dJdu_in, d_dJdu_in,
                               st);
                break;
            default: launch_hvp<8>(P, N, L, B, base, u_out, du_out, dJdu_out, d_dJdu_out, target, ldu, dJdu_in, d_dJdu_in,
                               st);
                break;
        }
    }
    if (hipGetLastError() != hipSuccess) return MESHPROP_ERR_HIP;
    // The partials are summed in workgroup order (the adjoint's reduction, accumulating).
    if (grad && (rc = launch_dp_reduce(L, N / 2, P.NP / 2, P.np, P.grid, P.E, (const float*)(base + P.off_pg), grad,
                                       ldp, 0, P.kso, st)))
        return rc;
    return launch_dp_reduce(L, N / 2, P.NP / 2, P.np, P.grid, P.E, (const float*)(base + P.off_ph), hvp, ldp, 0,
                            P.kso, st);
}

int meshprop_hvp(int X, int N, int L, int B, int fft, const int* geom0, const int* geom1, const float* p,
                 const float* q, int ldp, const float* s, int lds, const void* u_out, const void* du_out,
                 const void* dJdu_out, const void* d_dJdu_out, const void* target, float* grad, float* hvp, int ldu,
                 void* workspace, size_t workspace_bytes, void* stream) {
    return meshprop_hvp_grad_in(X, N, L, B, fft, geom0, geom1, p, q, ldp, s, lds, u_out, du_out, dJdu_out,
                                d_dJdu_out, target, grad, hvp, nullptr, nullptr, ldu, workspace, workspace_bytes,
                                stream);
}

}  // extern "C"
