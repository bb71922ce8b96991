// launch_fwddiff.hip -- instantiates and launches the fwddiff kernels.
#include "kernels.h"
#include "tiles.h"

namespace mp {

template <class T, int W, int V>
static int go(const Plan& P, const int* kind, const void* rows, const void* drows, const void* u_in,
              const void* du_in, void* u_out, void* du_out, int ldu, hipStream_t st) {
    typedef typename RowOf<T>::type R;
    if constexpr (W >= 8) {
        if (nn_fwddiff(P)) {
            // One vector per wave: one workgroup per NW vectors, each through the whole depth.
            constexpr int NW = kFwdDiffNNWaves;
            static std::atomic<unsigned long long> lds_set{0};
            const size_t shmem = nn_fwddiff_lds(P);
            if (int rc = lds_limit_once((const void*)k_fwddiff_nn<T, W>, lds_set, shmem)) return rc;
            const long grid = ((long)P.B + NW - 1) / NW;
            hipLaunchKernelGGL((k_fwddiff_nn<T, W>), dim3((unsigned)grid), dim3(64 * NW), shmem, st,
                               P.N, P.L, P.B, kind, (const R*)rows, (const R*)drows, (const T*)u_in, (const T*)du_in,
                               (T*)u_out, (T*)du_out, ldu);
            return MESHPROP_OK;
        }
    }
    hipLaunchKernelGGL((k_fwddiff<T, W, V>), dim3(P.grid), dim3(64 * P.waves), 0, st, P.N, P.L, P.B, P.LPV, P.NP,
                       kind, (const R*)rows, (const R*)drows, (const T*)u_in, (const T*)du_in, (T*)u_out, (T*)du_out,
                       ldu);
    return MESHPROP_OK;
}

int launch_fwddiff_any(const Plan& P, const int* kind, const void* rows, const void* drows, const void* u_in,
                       const void* du_in, void* u_out, void* du_out, int ldu, hipStream_t st) {
#define MP_CALL(W_, V_)                                                                        \
    if constexpr (diff_tile_ok(W_, V_)) {                                                      \
        if (P.real)                                                                            \
            return go<float, W_, V_>(P, kind, rows, drows, u_in, du_in, u_out, du_out, ldu, st); \
        return go<float2, W_, V_>(P, kind, rows, drows, u_in, du_in, u_out, du_out, ldu, st);    \
    }
    MP_FOR_WV(P.W, P.V, MP_CALL)
#undef MP_CALL
    return MESHPROP_ERR_UNSUPPORTED;
}

}  // namespace mp
