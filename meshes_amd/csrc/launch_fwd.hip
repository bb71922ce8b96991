// launch_fwd.hip -- instantiates and launches the fwdprop kernels (one translation unit
// per direction so the library builds in parallel).
#include "kernels.h"
#include "tiles.h"

namespace mp {

template <class T, int W, int V>
static void go(const Plan& P, const int* kind, const void* rows, int l0, int nl, const void* u_in, void* u_out,
               int ldu, hipStream_t st, Ckpt ck) {
    typedef typename RowOf<T>::type R;
    if constexpr (W == 16) {
        if (nn_forward(P, P.ncu)) {
            hipLaunchKernelGGL((k_fwd_nn<T, W, V>), dim3(P.grid), dim3(64 * P.waves), 0, st, P.N, l0, nl, P.B, kind,
                               (const R*)rows, (const T*)u_in, (T*)u_out, ldu, ck);
            return;
        }
    }
    hipLaunchKernelGGL((k_fwd<T, W, V>), dim3(P.grid), dim3(64 * P.waves), 0, st, P.N, l0, nl, P.B, P.LPV, P.NP, kind,
                       (const R*)rows, (const T*)u_in, (T*)u_out, ldu, ck);
}

#if MP_ILP_SET
int launch_fwd_ilp(const Plan& P, const int* kind, const void* rows, int l0, int nl, const void* u_in, void* u_out,
                   int ldu, hipStream_t st, Ckpt ck) {
#else
int launch_fwd_any(const Plan& P, const int* kind, const void* rows, int l0, int nl, const void* u_in, void* u_out,
                   int ldu, hipStream_t st, Ckpt ck) {
#endif
#define MP_CALL(W_, V_)                                                          \
    if constexpr (fwd_tile_ok(W_, V_)) {                                         \
        if constexpr (ilp_fwd_tile(W_, V_) == (MP_ILP_SET != 0)) {               \
            if (P.real)                                                          \
                go<float, W_, V_>(P, kind, rows, l0, nl, u_in, u_out, ldu, st, ck);  \
            else                                                                 \
                go<float2, W_, V_>(P, kind, rows, l0, nl, u_in, u_out, ldu, st, ck); \
            return MESHPROP_OK;                                                  \
        } else if constexpr (!MP_ILP_SET) {                                      \
            return launch_fwd_ilp(P, kind, rows, l0, nl, u_in, u_out, ldu, st, ck); \
        }                                                                        \
    }
    MP_FOR_WV(P.W, P.V, MP_CALL)
#undef MP_CALL
    return MESHPROP_ERR_UNSUPPORTED;
}

}  // namespace mp
