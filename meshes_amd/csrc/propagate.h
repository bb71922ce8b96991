// propagate.h -- register-resident layer kernels for gfx950 (wave64).
//
// Lane mapping.  A wave64 is split into G = 64/LPV lane groups; each group owns V field
// vectors and each of its LPV lanes holds W consecutive waveguides of every one of those
// vectors in VGPRs for the whole depth of the mesh ("field tile").  A layer is applied
// through per-waveguide *rows*: waveguide w of layer l becomes
//     u'_w = a_w u_w + b_w u_partner(w)
// where (a_w, b_w) is row w of the 2x2 crossing that acts on w (identity a=1, b=0 for
// waveguides no active crossing touches).  Rows are built once per call in FP64 by the
// prologue and streamed per layer from L2 (W rows per lane, coalesced), and every row load
// is reused for the V vectors the lane holds.
//
// Partners by layer kind (computed by the prologue, one code per layer):
//   0  NN, even shift : w ^ 1                      (in-lane)
//   1  NN, odd shift  : w+1 for odd w, w-1 for even w; the two ends of each lane's block
//                       pair with the neighbouring lanes (ds_bpermute)
//   2+st FFT, stride S = 2^st : w ^ S  (in-lane for S < W, lane ^ (S/W) otherwise)
//   255 identity layer (invalid stride): skipped
// These are the NN and FFT geometries of gpu/fwdprop.cu:48-112 (slot pairing of
// gpu/gmem.cu:236-240), restated for a row-per-waveguide layout.
#pragma once
#include <hip/hip_runtime.h>

namespace mp {

enum { KIND_IDENTITY = 255 };

// ---------------------------------------------------------------------------------------
// Scalar algebra: complex fields are float2 with rows float4 = (a.re, a.im, b.re, b.im);
// real fields are float with rows float2 = (a, b).  Every output costs 4 (complex) or 1
// (real) FMA-class instructions per term, i.e. 8 / 2 per waveguide and 16 / 4 per crossing.
// ---------------------------------------------------------------------------------------
template <class T>
struct RowOf;
template <>
struct RowOf<float2> {
    typedef float4 type;
};
template <>
struct RowOf<float> {
    typedef float2 type;
};

__device__ __forceinline__ float2 fzero(float2) { return make_float2(0.f, 0.f); }
__device__ __forceinline__ float fzero(float) { return 0.f; }

// a*x + b*y
__device__ __forceinline__ float2 rapply(const float4 r, const float2 x, const float2 y) {
    float2 o;
    o.x = fmaf(-r.w, y.y, fmaf(r.z, y.x, fmaf(-r.y, x.y, r.x * x.x)));
    o.y = fmaf(r.w, y.x, fmaf(r.z, y.y, fmaf(r.y, x.x, r.x * x.y)));
    return o;
}
__device__ __forceinline__ float rapply(const float2 r, const float x, const float y) {
    return fmaf(r.y, y, r.x * x);
}
// a*dx + b*dy + da*x + db*y
__device__ __forceinline__ float2 rapply_d(const float4 r, const float4 d, const float2 x, const float2 y,
                                           const float2 dx, const float2 dy) {
    float2 o;
    o.x = r.x * dx.x;
    o.x = fmaf(-r.y, dx.y, o.x);
    o.x = fmaf(r.z, dy.x, o.x);
    o.x = fmaf(-r.w, dy.y, o.x);
    o.x = fmaf(d.x, x.x, o.x);
    o.x = fmaf(-d.y, x.y, o.x);
    o.x = fmaf(d.z, y.x, o.x);
    o.x = fmaf(-d.w, y.y, o.x);
    o.y = r.x * dx.y;
    o.y = fmaf(r.y, dx.x, o.y);
    o.y = fmaf(r.z, dy.y, o.y);
    o.y = fmaf(r.w, dy.x, o.y);
    o.y = fmaf(d.x, x.y, o.y);
    o.y = fmaf(d.y, x.x, o.y);
    o.y = fmaf(d.z, y.y, o.y);
    o.y = fmaf(d.w, y.x, o.y);
    return o;
}
__device__ __forceinline__ float rapply_d(const float2 r, const float2 d, const float x, const float y,
                                          const float dx, const float dy) {
    return fmaf(d.y, y, fmaf(d.x, x, fmaf(r.y, dy, r.x * dx)));
}
// acc + conj(g) * x
__device__ __forceinline__ float2 cjfma(const float2 g, const float2 x, float2 acc) {
    acc.x = fmaf(g.y, x.y, fmaf(g.x, x.x, acc.x));
    acc.y = fmaf(-g.y, x.x, fmaf(g.x, x.y, acc.y));
    return acc;
}
__device__ __forceinline__ float cjfma(const float g, const float x, float acc) { return fmaf(g, x, acc); }
// acc.x + Im(conj(g) * x) (complex; .y is left alone); real fields have no imaginary part.
__device__ __forceinline__ float2 cjim(const float2 g, const float2 x, float2 acc) {
    acc.x = fmaf(-g.y, x.x, fmaf(g.x, x.y, acc.x));
    return acc;
}
__device__ __forceinline__ float cjim(const float, const float, const float acc) { return acc; }
__device__ __forceinline__ float2 fscl(const float2 a, const float s) { return make_float2(a.x * s, a.y * s); }
__device__ __forceinline__ float fscl(const float a, const float s) { return a * s; }

__device__ __forceinline__ float4 pack_row(const float2 a, const float2 b) { return make_float4(a.x, a.y, b.x, b.y); }
__device__ __forceinline__ float2 pack_row(const float a, const float b) { return make_float2(a, b); }
__device__ __forceinline__ float2 row_a(const float4 r) { return make_float2(r.x, r.y); }
__device__ __forceinline__ float2 row_b(const float4 r) { return make_float2(r.z, r.w); }
__device__ __forceinline__ float row_a(const float2 r) { return r.x; }
__device__ __forceinline__ float row_b(const float2 r) { return r.y; }
__device__ __forceinline__ float4 row_add(const float4 a, const float4 b) {
    return make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w);
}
__device__ __forceinline__ float2 row_add(const float2 a, const float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 fadd(const float2 a, const float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float fadd(const float a, const float b) { return a + b; }
__device__ __forceinline__ float2 fsub2(const float2 a, const float2 b) {
    return make_float2(2.f * (a.x - b.x), 2.f * (a.y - b.y));
}
__device__ __forceinline__ float fsub2(const float a, const float b) { return 2.f * (a - b); }

// Cross-lane read through the LDS crossbar (no LDS storage).  src is taken modulo 64.
__device__ __forceinline__ float lane_read(const float x, const int src) {
    return __int_as_float(__builtin_amdgcn_ds_bpermute(src << 2, __float_as_int(x)));
}
__device__ __forceinline__ float2 lane_read(const float2 x, const int src) {
    return make_float2(lane_read(x.x, src), lane_read(x.y, src));
}

// Neighbour reads through DPP (GFX9 wavefront shifts; no LDS round trip).  Lanes with no
// neighbour (lane 63 for next, lane 0 for prev) read 0; the rows there are the identity
// (b = 0), since they sit at the edge of the padded mesh.
__device__ __forceinline__ float lane_next(const float x) {  // value held by lane + 1
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x130, 0xf, 0xf, true));
}
__device__ __forceinline__ float lane_prev(const float x) {  // value held by lane - 1
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x138, 0xf, 0xf, true));
}
__device__ __forceinline__ float2 lane_next(const float2 x) { return make_float2(lane_next(x.x), lane_next(x.y)); }
__device__ __forceinline__ float2 lane_prev(const float2 x) { return make_float2(lane_prev(x.x), lane_prev(x.y)); }

template <int W>
struct Log2 {
    static constexpr int value = W == 1 ? 0 : W == 2 ? 1 : W == 4 ? 2 : W == 8 ? 3 : W == 16 ? 4 : -1;
};

// ---------------------------------------------------------------------------------------
// Layer kinds -> step variants.  Calls op.template run<KC, S>(xm) where KC = 0: in-lane
// pairs (k, k+S); KC = 1: odd-shift NN layer; KC = 2: cross-lane partner lane ^ xm.
// ---------------------------------------------------------------------------------------
template <int W, class Op>
__device__ __forceinline__ void dispatch_kind(const int kd, Op& op) {
    if (kd == 0 || kd == 2) {
        op.template run<0, 1>(0);
        return;
    }
    if (kd == 1) {
        op.template run<1, 1>(0);
        return;
    }
    if (kd == KIND_IDENTITY) {
        op.identity();
        return;
    }
    const int st = kd - 2;
    if constexpr (W > 2) {
        if (st == 1) {
            op.template run<0, 2>(0);
            return;
        }
    }
    if constexpr (W > 4) {
        if (st == 2) {
            op.template run<0, 4>(0);
            return;
        }
    }
    if constexpr (W > 8) {
        if (st == 3) {
            op.template run<0, 8>(0);
            return;
        }
    }
    op.template run<2, 1>(1 << (st - Log2<W>::value));
}

// ---------------------------------------------------------------------------------------
// Per-layer steps.  Odd-shift layers pair the last waveguide of each lane with the first
// of the next lane; those two values travel through DPP wavefront shifts.
// ---------------------------------------------------------------------------------------
template <class T, int W, int V>
__device__ __forceinline__ void copy_state(const T (&a)[V][W], T (&b)[V][W]) {
#pragma unroll
    for (int v = 0; v < V; ++v)
#pragma unroll
        for (int k = 0; k < W; ++k) b[v][k] = a[v][k];
}

// The state arrays come in pairs (in -> out): the kernels unroll the layer loop by two
// and alternate the roles, so every result is written to a register that is not an input
// of the same layer and no loop-carried copies are needed.
template <class T, int W, int V>
struct FwdStep {
    typedef typename RowOf<T>::type R;
    const T (&u)[V][W];
    T (&o)[V][W];
    const R (&r)[W];
    int lane;
    __device__ __forceinline__ void identity() { copy_state(u, o); }
    template <int KC, int S>
    __device__ __forceinline__ void run(const int xm) {
        if constexpr (KC == 0) {
#pragma unroll
            for (int v = 0; v < V; ++v)
#pragma unroll
                for (int k = 0; k < W; ++k)
                    if ((k & S) == 0) {
                        o[v][k] = rapply(r[k], u[v][k], u[v][k + S]);
                        o[v][k + S] = rapply(r[k + S], u[v][k + S], u[v][k]);
                    }
        } else if constexpr (KC == 1) {
            T hi[V], lo[V];
#pragma unroll
            for (int v = 0; v < V; ++v) {
                hi[v] = lane_next(u[v][0]);
                lo[v] = lane_prev(u[v][W - 1]);
            }
#pragma unroll
            for (int v = 0; v < V; ++v) {
#pragma unroll
                for (int k = 1; k + 1 < W; k += 2) {
                    o[v][k] = rapply(r[k], u[v][k], u[v][k + 1]);
                    o[v][k + 1] = rapply(r[k + 1], u[v][k + 1], u[v][k]);
                }
                o[v][W - 1] = rapply(r[W - 1], u[v][W - 1], hi[v]);
                o[v][0] = rapply(r[0], u[v][0], lo[v]);
            }
        } else {
            const int src = lane ^ xm;
#pragma unroll
            for (int v = 0; v < V; ++v)
#pragma unroll
                for (int k = 0; k < W; ++k) o[v][k] = rapply(r[k], u[v][k], lane_read(u[v][k], src));
        }
    }
};

template <class T, int W, int V>
struct FwdDiffStep {
    typedef typename RowOf<T>::type R;
    const T (&u)[V][W];
    const T (&du)[V][W];
    T (&o)[V][W];
    T (&dout)[V][W];
    const R (&r)[W];
    const R (&d)[W];
    int lane;
    __device__ __forceinline__ void identity() {
        copy_state(u, o);
        copy_state(du, dout);
    }
    __device__ __forceinline__ void one(const int v, const int k, const T y, const T dy) {
        dout[v][k] = rapply_d(r[k], d[k], u[v][k], y, du[v][k], dy);
        o[v][k] = rapply(r[k], u[v][k], y);
    }
    template <int KC, int S>
    __device__ __forceinline__ void run(const int xm) {
        if constexpr (KC == 0) {
#pragma unroll
            for (int v = 0; v < V; ++v)
#pragma unroll
                for (int k = 0; k < W; ++k) one(v, k, u[v][k ^ S], du[v][k ^ S]);
        } else if constexpr (KC == 1) {
            T hu[V], hd[V], lu[V], ld[V];
#pragma unroll
            for (int v = 0; v < V; ++v) {
                hu[v] = lane_next(u[v][0]);
                hd[v] = lane_next(du[v][0]);
                lu[v] = lane_prev(u[v][W - 1]);
                ld[v] = lane_prev(du[v][W - 1]);
            }
#pragma unroll
            for (int v = 0; v < V; ++v) {
#pragma unroll
                for (int k = 1; k + 1 < W; ++k) {
                    const int j = (k & 1) ? k + 1 : k - 1;
                    one(v, k, u[v][j], du[v][j]);
                }
                one(v, W - 1, hu[v], hd[v]);
                one(v, 0, lu[v], ld[v]);
            }
        } else {
            const int src = lane ^ xm;
#pragma unroll
            for (int v = 0; v < V; ++v)
#pragma unroll
                for (int k = 0; k < W; ++k) one(v, k, lane_read(u[v][k], src), lane_read(du[v][k], src));
        }
    }
};

// Reverse step through one layer: x = Rinv u (reconstructed layer input),
// dJdT rows A_k += conj(g_k) x_k, B_k += conj(g_k) x_partner(k), then g <- Radj g.
// KF (unitary crossings): the rows are taken against the layer-input adjoint g' = T^H g
// instead, and only what dp needs is kept: A_k.x += Im(conj(g'_k) x_k) (A_k.y stays 0; real
// fields keep no A) and B_k += conj(g'_k) x_partner(k).  With g = T g' the gradient is
// dJ/dp = Re sum_rc K_rc Q_rc, Q_rc = sum_b conj(g'_r) x_c, K = T^H dT/dp anti-Hermitian,
// which needs Im Q00, Im Q11 and Q01 - conj(Q10) only: 12 FMA per crossing and vector
// instead of 16.  SC (scaled-unitary crossings T = c U): the adjoint rows are |c|^2 times the
// inverse rows, g' = sc_k (Rinv g), with the per-waveguide factor sc of the layer.
template <class T, int W, int V, bool KF = false, bool SC = false>
struct BwdStep {
    typedef typename RowOf<T>::type R;
    const T (&u)[V][W];
    const T (&g)[V][W];
    T (&uo)[V][W];
    T (&go)[V][W];
    const R (&ri)[W];
    const R (&ra)[W];
    T (&accA)[W];
    T (&accB)[W];
    int lane;
    const float* sc;  // SC: this layer's W factors (registers after unrolling)
    __device__ __forceinline__ T adj(const int k, const T gk, const T gp) const {
        const T r = rapply(ra[k], gk, gp);
        if constexpr (SC)
            return fscl(r, sc[k]);
        else
            return r;
    }
    __device__ __forceinline__ void identity() {
        // Identity layer: fields pass through; its rows meet a zero dT/dp.
        copy_state(u, uo);
        copy_state(g, go);
    }
    // Accumulate waveguide k's dJdT row from its reconstructed input xk, its partner's xp and
    // (KF) its layer-input adjoint gin, else its layer-output adjoint gout.
    __device__ __forceinline__ void acc(const int k, const T gout, const T gin, const T xk, const T xp) {
        if constexpr (KF) {
            accA[k] = cjim(gin, xk, accA[k]);
            accB[k] = cjfma(gin, xp, accB[k]);
        } else {
            accA[k] = cjfma(gout, xk, accA[k]);
            accB[k] = cjfma(gout, xp, accB[k]);
        }
    }
    __device__ __forceinline__ void pair(const int v, const int k, const int j) {
        const T xk = rapply(ri[k], u[v][k], u[v][j]), xj = rapply(ri[j], u[v][j], u[v][k]);
        const T gk = adj(k, g[v][k], g[v][j]), gj = adj(j, g[v][j], g[v][k]);
        acc(k, g[v][k], gk, xk, xj);
        acc(j, g[v][j], gj, xj, xk);
        go[v][k] = gk;
        go[v][j] = gj;
        uo[v][k] = xk;
        uo[v][j] = xj;
    }
    template <int KC, int S>
    __device__ __forceinline__ void run(const int xm) {
        if constexpr (KC == 0) {
#pragma unroll
            for (int v = 0; v < V; ++v)
#pragma unroll
                for (int k = 0; k < W; ++k)
                    if ((k & S) == 0) pair(v, k, k + S);
        } else if constexpr (KC == 1) {
#pragma unroll
            for (int v = 0; v < V; ++v) {
                const T hu = lane_next(u[v][0]), hg = lane_next(g[v][0]);
                const T lu = lane_prev(u[v][W - 1]), lg = lane_prev(g[v][W - 1]);
                uo[v][W - 1] = rapply(ri[W - 1], u[v][W - 1], hu);
                uo[v][0] = rapply(ri[0], u[v][0], lu);
                go[v][W - 1] = adj(W - 1, g[v][W - 1], hg);
                go[v][0] = adj(0, g[v][0], lg);
            }
#pragma unroll
            for (int v = 0; v < V; ++v) {
                const T hx = lane_next(uo[v][0]), lx = lane_prev(uo[v][W - 1]);
#pragma unroll
                for (int k = 1; k + 1 < W; k += 2) pair(v, k, k + 1);
                acc(W - 1, g[v][W - 1], go[v][W - 1], uo[v][W - 1], hx);
                acc(0, g[v][0], go[v][0], uo[v][0], lx);
            }
        } else {
            const int src = lane ^ xm;
#pragma unroll
            for (int v = 0; v < V; ++v)
#pragma unroll
                for (int k = 0; k < W; ++k) {
                    uo[v][k] = rapply(ri[k], u[v][k], lane_read(u[v][k], src));
                    go[v][k] = adj(k, g[v][k], lane_read(g[v][k], src));
                }
#pragma unroll
            for (int v = 0; v < V; ++v)
#pragma unroll
                for (int k = 0; k < W; ++k) acc(k, g[v][k], go[v][k], uo[v][k], lane_read(uo[v][k], src));
        }
    }
};

}  // namespace mp
