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
