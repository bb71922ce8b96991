// launch_bwd.hip -- instantiates and launches the backdiff kernels (unitary and general).
#include "kernels.h"
#include "tiles.h"

namespace mp {

template <class T, int W, int V>
static void go(const Plan& P, const int* kind, const void* ri, const void* ra, const void* dtdp, float* partial,
               size_t ps, int l0, int nl, const void* u_out, const void* g_out, const void* target, void* u_in,
               void* g_in, int ldu, hipStream_t st, Ckpt ck) {
    typedef typename RowOf<T>::type R;
    const size_t shmem = (size_t)P.waves * P.C * P.NP * P.rowbytes;
    if (shmem > 64 * 1024) {  // beyond the default dynamic LDS limit of a launch
        if (P.uni)
            hipFuncSetAttribute((const void*)k_bwd<T, W, V, true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)shmem);
        else
            hipFuncSetAttribute((const void*)k_bwd<T, W, V, false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)shmem);
    }
    if (P.uni)
        hipLaunchKernelGGL((k_bwd<T, W, V, true>), dim3(P.grid), dim3(64 * P.waves), shmem, st, P.N, l0, nl, P.B,
                           P.LPV, P.NP, P.C, P.np, kind, (const R*)ri, (const R*)ra, (const T*)dtdp, partial, ps,
                           P.ntiles, (const T*)u_out, (const T*)g_out, (const T*)target, (T*)u_in, (T*)g_in, ldu, ck);
    else
        hipLaunchKernelGGL((k_bwd<T, W, V, false>), dim3(P.grid), dim3(64 * P.waves), shmem, st, P.N, l0, nl, P.B,
                           P.LPV, P.NP, P.C, P.np, kind, (const R*)ri, (const R*)ra, (const T*)dtdp, partial, ps,
                           P.ntiles, (const T*)u_out, (const T*)g_out, (const T*)target, (T*)u_in, (T*)g_in, ldu, ck);
}

int launch_bwd_any(const Plan& P, const int* kind, const void* ri, const void* ra, const void* dtdp, float* partial,
                   size_t ps, int l0, int nl, const void* u_out, const void* g_out, const void* target, void* u_in,
                   void* g_in, int ldu, hipStream_t st, Ckpt ck) {
#define MP_CALL(W_, V_)                                                                                         \
    if constexpr (diff_tile_ok(W_, V_)) {                                                                       \
        if (P.real)                                                                                             \
            go<float, W_, V_>(P, kind, ri, ra, dtdp, partial, ps, l0, nl, u_out, g_out, target, u_in, g_in, ldu, \
                              st, ck);                                                                              \
        else                                                                                                    \
            go<float2, W_, V_>(P, kind, ri, ra, dtdp, partial, ps, l0, nl, u_out, g_out, target, u_in, g_in,    \
                               ldu, st, ck);                                                                        \
        return MESHPROP_OK;                                                                                     \
    }
    MP_FOR_WV(P.W, P.V, MP_CALL)
#undef MP_CALL
    return MESHPROP_ERR_UNSUPPORTED;
}

}  // namespace mp
