// launch_fwd.hip -- instantiates and launches the fwdprop kernels (one translation unit
// per direction so the library builds in parallel).
#include "kernels.h"
#include "tiles.h"

namespace mp {

template <class T, int W, int V>
static void go(const Plan& P, const int* kind, const void* rows, const void* u_in, void* u_out, int ldu, void* ckpt,
               int ck, hipStream_t st) {
    typedef typename RowOf<T>::type R;
    hipLaunchKernelGGL((k_fwd<T, W, V>), dim3(P.grid), dim3(64 * P.waves), 0, st, P.N, P.L, P.B, P.LPV, P.NP, kind,
                       (const R*)rows, (const T*)u_in, (T*)u_out, ldu, (T*)ckpt, ck);
}

int launch_fwd_any(const Plan& P, const int* kind, const void* rows, const void* u_in, void* u_out, int ldu,
                   void* ckpt, int ck, hipStream_t st) {
#define MP_CALL(W_, V_)                                                          \
    if constexpr (fwd_tile_ok(W_, V_)) {                                         \
        if (P.real)                                                              \
            go<float, W_, V_>(P, kind, rows, u_in, u_out, ldu, ckpt, ck, st);    \
        else                                                                     \
            go<float2, W_, V_>(P, kind, rows, u_in, u_out, ldu, ckpt, ck, st);   \
        return MESHPROP_OK;                                                      \
    }
    MP_FOR_WV(P.W, P.V, MP_CALL)
#undef MP_CALL
    return MESHPROP_ERR_UNSUPPORTED;
}

}  // namespace mp
