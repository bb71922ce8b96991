// tiles.h -- run-time dispatch over the compiled (W, V) tiles.
#pragma once

#define MP_FOR_WV(W_, V_, CALL)              \
    switch ((W_) * 16 + (V_)) {              \
        case 2 * 16 + 1: { CALL(2, 1) } break;   \
        case 2 * 16 + 2: { CALL(2, 2) } break;   \
        case 2 * 16 + 4: { CALL(2, 4) } break;   \
        case 2 * 16 + 8: { CALL(2, 8) } break;   \
        case 4 * 16 + 1: { CALL(4, 1) } break;   \
        case 4 * 16 + 2: { CALL(4, 2) } break;   \
        case 4 * 16 + 4: { CALL(4, 4) } break;   \
        case 4 * 16 + 8: { CALL(4, 8) } break;   \
        case 8 * 16 + 1: { CALL(8, 1) } break;   \
        case 8 * 16 + 2: { CALL(8, 2) } break;   \
        case 8 * 16 + 4: { CALL(8, 4) } break;   \
        case 8 * 16 + 8: { CALL(8, 8) } break;   \
        case 16 * 16 + 1: { CALL(16, 1) } break; \
        case 16 * 16 + 2: { CALL(16, 2) } break; \
        case 16 * 16 + 4: { CALL(16, 4) } break; \
        case 16 * 16 + 8: { CALL(16, 8) } break; \
        default: break;                          \
    }

// Tiles whose fwdprop and backdiff kernels are compiled with LLVM's ILP-oriented machine
// schedulers (launch_fwd_ilp.hip and launch_bwd_ilp.hip: iterative-ilp; the flag is per
// translation unit).  Measured at N=256, L=256, B=4096 (two waves per SIMD): 85.6 / 320 us for
// fwdprop / backdiff against 89.2 / 329 us with the default scheduler; with the later adjoint,
// iterative-ilp 272 us against 276 us for max-ilp.  Applied to every tile it raised register use
// elsewhere: spills in the wide lossy adjoint and lower occupancy for the small tiles.
constexpr bool ilp_tile(int W, int V) { return W == 4 && V == 2; }
// The forward's set adds the W = 16 tiles: N=1024 L=1024 B=2048 fwdprop 1.07 ms against 1.14 ms
// with the default scheduler (the W = 16 adjoint is slower under it and stays out).
constexpr bool ilp_fwd_tile(int W, int V) { return ilp_tile(W, V) || W == 16; }

// Each launch_*.hip file is compiled twice: MP_ILP_SET = 0 holds the other tiles and forwards
// the ILP set to the second object, MP_ILP_SET = 1 (the *_ilp.hip wrapper) holds the ILP set.
#ifndef MP_ILP_SET
#define MP_ILP_SET 0
#endif
