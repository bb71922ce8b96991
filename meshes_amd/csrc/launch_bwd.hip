// launch_bwd.hip -- instantiates and launches the backdiff kernels (unitary and general).
// Compiled three times: the default object (MP_ILP_SET = MP_SC_SET = 0) holds every tile
// except ilp_tile() and forwards those to launch_bwd_ilp.hip (MP_ILP_SET = 1, iterative-ilp
// scheduler); launch_bwd_sc.hip (MP_SC_SET = 1) holds the scaled-unitary adjoints of every
// complex tile.
#include "kernels.h"
#include "tiles.h"

#ifndef MP_SC_SET
#define MP_SC_SET 0
#endif

namespace mp {

template <class T, int W, int V, bool UNI, int PIPE, bool SC>
static int go1(const Plan& P, const int* kind, const void* ri, const void* ra, const void* dtdp, float* partial,
               size_t ps, int l0, int nl, const void* u_out, const void* g_out, const void* target, void* u_in,
               void* g_in, int ldu, hipStream_t st, Ckpt ck) {
    typedef typename RowOf<T>::type R;
    const size_t shmem = (size_t)P.waves * P.C * stage_layer_bytes(P) * (PIPE ? 2 : 1);
    // W = 16 unitary adjoints: the entry compiled for two waves per SIMD.
    constexpr bool w16 = W >= 16 && UNI;
    static std::atomic<unsigned long long> lds_set{0};  // devices whose LDS limit this kernel has raised
    const void* fn;
    if constexpr (w16)
        fn = (const void*)k_bwd_w16<T, W, V, UNI, PIPE, SC>;
    else
        fn = (const void*)k_bwd<T, W, V, UNI, PIPE, SC>;
    if (int rc = lds_limit_once(fn, lds_set, shmem)) return rc;
    if constexpr (w16)
        hipLaunchKernelGGL((k_bwd_w16<T, W, V, UNI, PIPE, SC>), dim3(P.grid), dim3(64 * P.waves), shmem, st, P.N, l0,
                           nl, P.B, P.LPV, P.NP, P.C, P.np, kind, (const R*)ri, (const R*)ra, (const T*)dtdp, partial,
                           ps, P.ntiles, (const T*)u_out, (const T*)g_out, (const T*)target, (T*)u_in, (T*)g_in, ldu,
                           ck, P.scales);
    else
        hipLaunchKernelGGL((k_bwd<T, W, V, UNI, PIPE, SC>), dim3(P.grid), dim3(64 * P.waves), shmem, st, P.N, l0, nl,
                           P.B, P.LPV, P.NP, P.C, P.np, kind, (const R*)ri, (const R*)ra, (const T*)dtdp, partial, ps,
                           P.ntiles, (const T*)u_out, (const T*)g_out, (const T*)target, (T*)u_in, (T*)g_in, ldu, ck,
                           P.scales);
    return MESHPROP_OK;
}

// NN-specialised unitary adjoint (k_bwd_nn) for NW-wave workgroups; scaled-unitary crossings
// in the scaled object.
template <class T, int W, int V, int NW, bool RS = nn_rows_staged(W, NW, MP_SC_SET != 0)>
static int go_nn(const Plan& P, const int* kind, const void* ri, const void* dtdp, float* partial, size_t ps, int l0,
                 int nl, const void* u_out, const void* g_out, const void* target, void* u_in, void* g_in, int ldu,
                 hipStream_t st, Ckpt ck) {
    typedef typename RowOf<T>::type R;
    constexpr int C = nn_stage_depth(W);
    constexpr bool SC = MP_SC_SET != 0;
    if constexpr (W >= 16 && NW == 4 && !SC && !RS) {
        if (nn_rows_staged_small(P))
            return go_nn<T, W, V, NW, true>(P, kind, ri, dtdp, partial, ps, l0, nl, u_out, g_out, target, u_in, g_in,
                                            ldu, st, ck);
    }
    const size_t shmem = nn_lds_bytes(W, NW, C, RS, sizeof(R));
    static std::atomic<unsigned long long> lds_set{0};
    // W = 16: the entry compiled for two waves per SIMD.
    auto kern = [] {
        if constexpr (W >= 16)
            return k_bwd_nn_w16<T, W, V, NW, C, SC, RS>;
        else
            return k_bwd_nn<T, W, V, NW, C, SC, RS>;
    }();
    if (int rc = lds_limit_once((const void*)kern, lds_set, shmem)) return rc;
    hipLaunchKernelGGL(kern, dim3(P.grid), dim3(64 * NW), shmem, st, P.N, l0, nl, P.B, kind, (const R*)ri,
                       (const T*)dtdp, partial, ps, P.ntiles, (const T*)u_out, (const T*)g_out, (const T*)target,
                       (T*)u_in, (T*)g_in, ldu, ck, P.scales);
    return MESHPROP_OK;
}

#define MP_GO_ARGS P, kind, ri, ra, dtdp, partial, ps, l0, nl, u_out, g_out, target, u_in, g_in, ldu, st, ck
template <class T, int W, int V, int PIPE>
static int go2(const Plan& P, const int* kind, const void* ri, const void* ra, const void* dtdp, float* partial,
               size_t ps, int l0, int nl, const void* u_out, const void* g_out, const void* target, void* u_in,
               void* g_in, int ldu, hipStream_t st, Ckpt ck) {
    if constexpr (MP_SC_SET) {
        // Scaled unitaries are complex (generic MZI) and run the unitary adjoint.
        if constexpr (std::is_same<T, float2>::value) return go1<T, W, V, true, PIPE, true>(MP_GO_ARGS);
    } else if (P.uni) {
        return go1<T, W, V, true, PIPE, false>(MP_GO_ARGS);
    } else if constexpr (PIPE == 0 || W <= 4) {  // the general stage fits two buffers only for W <= 4
        return go1<T, W, V, false, PIPE, false>(MP_GO_ARGS);
    }
    return MESHPROP_ERR_UNSUPPORTED;
}

template <class T, int W, int V>
static int go(const Plan& P, const int* kind, const void* ri, const void* ra, const void* dtdp, float* partial,
              size_t ps, int l0, int nl, const void* u_out, const void* g_out, const void* target, void* u_in,
              void* g_in, int ldu, hipStream_t st, Ckpt ck) {
    // Unitary (or, in the scaled object, scaled-unitary) adjoints on NN geometries.
    if constexpr (std::is_same<T, float2>::value && W >= 2 && (W < 16 || V == 1)) {
        if (nn_adjoint(P) && (MP_SC_SET || !P.scaled)) {
            if (P.waves == 8)
                return go_nn<T, W, V, 8>(P, kind, ri, dtdp, partial, ps, l0, nl, u_out, g_out, target, u_in, g_in, ldu,
                                         st, ck);
            return go_nn<T, W, V, 4>(P, kind, ri, dtdp, partial, ps, l0, nl, u_out, g_out, target, u_in, g_in, ldu,
                                     st, ck);
        }
    }
    // The general pipelined stage is compiled for the tiles it can serve (two buffers of two
    // layers fit the LDS budget for W <= 4, and for W = 8 with the per-crossing K-form stage),
    // once for four-wave and once for eight-wave groups.
    if constexpr (W <= 8) {
        if (P.pipe) {
            if (P.waves > 4) return go2<T, W, V, kPipeWaves>(MP_GO_ARGS);
            return go2<T, W, V, 4>(MP_GO_ARGS);
        }
    }
    return go2<T, W, V, 0>(MP_GO_ARGS);
}

// This object's tiles: every tile for the scaled set, else the ILP split.
constexpr bool own_tile(int W, int V) { return MP_SC_SET || ilp_tile(W, V) == (MP_ILP_SET != 0); }

#if MP_SC_SET
int launch_bwd_sc(const Plan& P, const int* kind, const void* ri, const void* ra, const void* dtdp, float* partial,
                  size_t ps, int l0, int nl, const void* u_out, const void* g_out, const void* target, void* u_in,
                  void* g_in, int ldu, hipStream_t st, Ckpt ck) {
#elif MP_ILP_SET
int launch_bwd_ilp(const Plan& P, const int* kind, const void* ri, const void* ra, const void* dtdp, float* partial,
                   size_t ps, int l0, int nl, const void* u_out, const void* g_out, const void* target, void* u_in,
                   void* g_in, int ldu, hipStream_t st, Ckpt ck) {
#else
int launch_bwd_any(const Plan& P, const int* kind, const void* ri, const void* ra, const void* dtdp, float* partial,
                   size_t ps, int l0, int nl, const void* u_out, const void* g_out, const void* target, void* u_in,
                   void* g_in, int ldu, hipStream_t st, Ckpt ck) {
    if (P.scaled)
        return P.real ? MESHPROP_ERR_UNSUPPORTED
                      : launch_bwd_sc(P, kind, ri, ra, dtdp, partial, ps, l0, nl, u_out, g_out, target, u_in, g_in,
                                      ldu, st, ck);
#endif
#define MP_CALL(W_, V_)                                                                                           \
    if constexpr (diff_tile_ok(W_, V_)) {                                                                         \
        if constexpr (own_tile(W_, V_)) {                                                                         \
            if (P.real) return go<float, W_, V_>(MP_GO_ARGS);                                                    \
            return go<float2, W_, V_>(MP_GO_ARGS);                                                                \
        } else if constexpr (!MP_ILP_SET && !MP_SC_SET) {                                                         \
            return launch_bwd_ilp(P, kind, ri, ra, dtdp, partial, ps, l0, nl, u_out, g_out, target, u_in, g_in,   \
                                  ldu, st, ck);                                                                   \
        }                                                                                                         \
    }
    MP_FOR_WV(P.W, P.V, MP_CALL)
#undef MP_CALL
#undef MP_GO_ARGS
    return MESHPROP_ERR_UNSUPPORTED;
}

}  // namespace mp
