// meshprop.hip -- libmeshprop.so: MI355X (gfx950) batched propagation through meshes of
// 2x2 crossings, and its C ABI (include/meshprop.h).
//
// One call = (1) a prologue kernel that turns slot-layout parameters and the layer
// geometry into per-waveguide row tables (FP64 maths, FP32 storage) plus one kind code
// per layer, (2) one propagation kernel that keeps each field tile in VGPRs across all L
// layers, and (3) for backdiff, a fixed-order reduction of per-workgroup dp partials.
// Kernels never allocate: everything lives in the caller's workspace (or, for
// meshprop_run_noalloc, in the library's static device scratch).
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <mutex>

#include "crossing.h"
#include "meshprop.h"
#include "propagate.h"

#include "kernels.h"

namespace mp {

// dp reduction: dp[l*ldp + m*np + k] = (or +=) the sum over workgroups w of
// partial[w][(l*NPh + m)*np + k], m < N/2.  A block is four waves over 64 consecutive dp
// entries; wave q sums workgroups [q*nwg/4, (q+1)*nwg/4) with four interleaved accumulators and
// the four wave sums are added in wave order through LDS: one launch, a fixed summation order
// (bitwise reproducible), and four waves' worth of loads in flight per entry.
// kso > 0: the partials are in the NN adjoint's stage order (position p*64 + t holds slot
// t*kso + p); the blocks then run over positions (coalesced partial reads) and scatter to dp.
__global__ void __launch_bounds__(256) k_dp_reduce(int L, int Nh, int NPh, int np, int nwg, size_t E,
                                                   const float* __restrict__ partial, float* __restrict__ dp, int ldp,
                                                   int assign, int kso) {
    __shared__ float red[4][64];
    const int lane = threadIdx.x & 63, wq = threadIdx.x >> 6;
    const long idx = (long)blockIdx.x * 64 + lane;
    const int Nr = kso > 0 ? NPh : Nh;  // entries per layer the blocks run over
    const long total = (long)L * Nr * np;
    bool ok = idx < total;
    int k = 0, mm = 0, l = 0, pos = 0;
    if (ok) {
        k = (int)(idx % np);
        const long lm = idx / np;
        pos = (int)(lm % Nr);
        l = (int)(lm / Nr);
        mm = stage_slot(pos, kso);
        ok = mm < Nh;
    }
    float acc = 0.f;
    if (ok) {
        const float* src = partial + ((size_t)l * NPh + pos) * np + k;
        const int w0 = (int)((long)nwg * wq / 4), w1 = (int)((long)nwg * (wq + 1) / 4);
        float a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f;
        int w = w0;
        for (; w + 3 < w1; w += 4) {
            a0 += src[(size_t)w * E];
            a1 += src[(size_t)(w + 1) * E];
            a2 += src[(size_t)(w + 2) * E];
            a3 += src[(size_t)(w + 3) * E];
        }
        for (; w < w1; ++w) a0 += src[(size_t)w * E];
        acc = (a0 + a1) + (a2 + a3);
    }
    red[wq][lane] = acc;
    __syncthreads();
    if (wq == 0 && ok) {
        const float sum = ((red[0][lane] + red[1][lane]) + red[2][lane]) + red[3][lane];
        float* d = dp + (size_t)l * ldp + (size_t)mm * np + k;
        *d = assign ? sum : *d + sum;
    }
}

// Host launcher of k_dp_reduce for the other translation units (the HVP reductions):
// dp[l*ldp + m*np + k] (+)= the nwg partials (stride E) of slot m, in workgroup order.
int launch_dp_reduce(int L, int Nh, int NPh, int np, int nwg, size_t E, const float* partial, float* dp, int ldp,
                     int assign, int kso, hipStream_t st) {
    const long total = (long)L * (kso > 0 ? NPh : Nh) * np;
    if (total <= 0) return MESHPROP_OK;
    hipLaunchKernelGGL(k_dp_reduce, dim3((unsigned)((total + 63) / 64)), dim3(256), 0, st, L, Nh, NPh, np, nwg, E,
                       partial, dp, ldp, assign, kso);
    return hipGetLastError() == hipSuccess ? MESHPROP_OK : MESHPROP_ERR_HIP;
}

}  // namespace mp


// =======================================================================================
// Host side
// =======================================================================================
using namespace mp;

namespace {

thread_local int g_last_hip_error = 0;

inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// LDS of the backdiff stage.  gfx950 gives a workgroup up to 160 KiB; workgroups of up to
// four waves keep to 64 KiB so that two of them share a CU.
constexpr size_t kStageBytes = 64 * 1024;
constexpr size_t kStageBytesWide = 128 * 1024;
inline size_t stage_budget(int waves) { return waves > 4 ? kStageBytesWide : kStageBytes; }

int num_cus() {
    int dev = 0, n = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 256;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) return 256;
    return n;
}

int pow2ceil(int x) {
    int p = 1;
    while (p < x) p <<= 1;
    return p;
}

bool valid_w(int w) { return w == 2 || w == 4 || w == 8 || w == 16; }

bool tile_ok(int op, int W, int V) {
    return op == MESHPROP_OP_FWDPROP ? fwd_tile_ok(W, V) : diff_tile_ok(W, V);
}

// Measured tile table of the plain path (gpu/__init__.py:26-35 keeps the reference's measured
// warps per kernel, crossing type, square/fat batch and N bucket): the fastest tile of whole
// meshprop_run calls on square N x N x N and fat N x N x 4096 Clements meshes, written by
// meshes_amd/tune_tiles.py.  A call is fat when B > 2N (gpu/mesh.py:161) and takes the entry of
// the next tabulated N at or above its own (gpu/mesh.py:84).
struct TileEntry {
    int op, X, fat, n0, w, v, waves;
};
const TileEntry kTileTable[] = {
#include "tile_table.inc"
    {-1, -1, 0, 0, 0, 0, 0},  // end (matches no call)
};

const TileEntry* table_tile(int op, int X, int N, int B) {
    const int fat = (long)B > 2L * N ? 1 : 0;
    const TileEntry* best = nullptr;
    for (const TileEntry& e : kTileTable)
        if (e.op == op && e.X == X && e.fat == fat && e.n0 >= N && (!best || e.n0 < best->n0)) best = &e;
    return best;
}

// Default tile: narrowest W that keeps a vector inside one wave; the largest compiled V
// that still gives two waves per SIMD (8 per CU).
void resolve_tile(int op, int N, int B, const meshprop_tile* t, int ncu, int& W, int& V, int& LPV, int& waves,
                  int& C) {
    W = N <= 128 ? 2 : N <= 256 ? 4 : N <= 512 ? 8 : 16;
    if (t && valid_w(t->w) && (long)t->w * 64 >= N) W = t->w;
    LPV = pow2ceil((N + W - 1) / W);
    const int G = 64 / LPV;
    waves = (t && t->waves >= 1 && t->waves <= 8) ? t->waves : 4;
    if (t && t->v >= 1 && tile_ok(op, W, t->v) && (t->v & (t->v - 1)) == 0) {
        V = t->v;
    } else {
        V = 1;
        const long target = 8L * ncu;  // two waves per SIMD hide more latency than wider tiles
        for (int v = 8; v >= 1; v >>= 1) {
            if (!tile_ok(op, W, v)) continue;
            const long wv = ((long)B + (long)G * v - 1) / ((long)G * v);
            if (wv >= target) {
                V = v;
                break;
            }
        }
    }
    C = (t && t->chunk >= 1) ? t->chunk : 0;
}

// grid_cap > 0 limits the adjoint's workgroups (and with them its dp partials) further; the
// workgroups then walk more vector tiles each.
int make_plan(int op, int X, int N, int L, int B, int fft, const meshprop_tile* tile, Plan& P, int flags = 0,
              int grid_cap = 0) {
    if (X < 0 || X > 3) return MESHPROP_ERR_ARG;
    if (op < 0 || op > 3) return MESHPROP_ERR_ARG;
    const bool adj = (op == MESHPROP_OP_BACKDIFF || op == MESHPROP_OP_FWDGRAD);
    // Generic crossings asserted to be scaled unitaries run the unitary adjoint.
    P.scaled = adj && X == X_GENERIC && (flags & MESHPROP_SCALED_UNITARY) != 0;
    P.scales = nullptr;
    P.sorder = 0;
    P.nn = (flags & MESHPROP_GENERIC_KERNELS) ? 0 : 1;
    // The fused forward + adjoint runs the unitary adjoint: unitary or scaled-unitary crossings.
    if (op == MESHPROP_OP_FWDGRAD && X == X_GENERIC && !P.scaled) return MESHPROP_ERR_UNSUPPORTED;
    if (N <= 0 || (N & 1) || L < 0 || B < 0) return MESHPROP_ERR_ARG;
    if (N > 1024) return MESHPROP_ERR_UNSUPPORTED;
    P.op = op;
    P.X = X;
    P.N = N;
    P.L = L;
    P.B = B;
    P.fft = fft;
    P.np = n_phase_of(X);
    P.ns = n_split_of(X);
    P.real = (X == X_ORTH);
    P.uni = (X != X_GENERIC) || P.scaled;
    const int ncu = num_cus();
    P.ncu = ncu;
    // Without an explicit tile, neighbour meshes take the measured table's (a requested stage
    // depth is kept); butterfly meshes and untabulated cases take the heuristic.
    meshprop_tile tabled;
    const bool explicit_tile = tile && (tile->w > 0 || tile->v > 0 || tile->waves > 0);
    if (!explicit_tile && !fft) {
        // Scaled-unitary crossings run the unitary adjoint: they take the MZI entries.
        if (const TileEntry* e = table_tile(op, P.scaled ? X_MZI : X, N, B)) {
            if ((long)e->w * 64 >= N && tile_ok(op, e->w, e->v)) {
                tabled = meshprop_tile{e->w, e->v, e->waves, tile ? tile->chunk : 0};
                // W = 16 unitary adjoints pick their workgroup size by batch (below).
                if (adj && P.uni && e->w == 16) tabled.waves = 0;
                tile = &tabled;
            }
        }
    }
    resolve_tile(op, N, B, tile, ncu, P.W, P.V, P.LPV, P.waves, P.C);
    // Unitary adjoints at W = 8 hold one vector per lane unless asked otherwise: the two-vector
    // tile needs one wave per SIMD (pipelined, eight waves: spills).  Measured at N=512, L=512,
    // B=8192 with checkpoints every 32 layers: 3.52 ms (8,1) pipelined against 4.10 ms (8,2).
    if (adj && P.uni && P.W == 8 && !(tile && tile->v >= 1)) P.V = 1;
    P.NP = P.LPV * P.W;
    P.rowbytes = P.real ? sizeof(float2) : sizeof(float4);
    // Widest unitary NN adjoint (W = 16, one vector per lane): the NN-specialised kernel with a
    // pipelined stage of one layer (two buffers of 8 KiB per wave).
    const bool nn16 = adj && P.nn && P.W == 16 && P.V == 1 && P.uni && !P.real && !fft && P.LPV == 64 &&
                      !(tile && tile->chunk > 1);
    {
        // Eight-wave adjoint workgroups exist only with the pipelined stage (two buffers of two
        // layers of eight waves in the wide budget; one layer at W = 16).
        const bool pipe_w = P.W <= 4 || (P.W == 8 && P.uni);  // widths with a pipelined stage
        const bool wide_ok = adj && ((pipe_w && (P.NP / 2) % 64 == 0 &&
                                      (size_t)4 * kPipeWaves * stage_layer_bytes(P) <= stage_budget(kPipeWaves)) ||
                                     nn16);
        const int cap = adj ? (wide_ok ? kPipeWaves : bwd_max_waves(P.W, P.V, P.uni)) : 4;
        // Automatic choice: eight-wave adjoint workgroups for unitary crossings when the batch
        // still fills every CU with them (half the dp partials of four-wave groups; measured at
        // N=256, L=256, B=4096: 0.406-0.413 ms per fwd+grad step against 0.419 ms).  The
        // general (lossy) adjoint keeps four waves: its eight-wave kernel spills registers.
        // The NN-specialised W = 16 adjoint takes eight waves already when four-wave groups would
        // fill every CU (MI355X, N=1024 L=1024 fwd+grad: B=1024 3.145 ms with eight waves against
        // 3.486 with four; B=16..256 2.94-2.95 against 2.83-2.86; B=2048 backdiff 3.21 against 3.41).
        const bool auto_waves = !(tile && tile->waves >= 1 && tile->waves <= 8);
        if (auto_waves && wide_ok && P.uni) {
            const long per_wg = (long)(nn16 ? 4 : kPipeWaves) * (64 / P.LPV) * P.V;
            if (((long)B + per_wg - 1) / per_wg >= ncu) P.waves = kPipeWaves;
        }
        if (P.waves > cap) P.waves = cap;
        if (adj && P.waves > 4 && P.waves < kPipeWaves) P.waves = 4;
    }
    if (adj)  // the LDS stage must hold one unrolled group of layers
        while (P.waves > 1 && (size_t)(P.W <= 4 ? 4 : 2) * P.waves * stage_layer_bytes(P) > stage_budget(P.waves))
            P.waves >>= 1;
    const int G = 64 / P.LPV;
    const long per_tile = (long)P.waves * G * P.V;
    P.ntiles = ((long)B + per_tile - 1) / per_tile;
    if (P.ntiles < 1) P.ntiles = 1;
    P.grid = (int)P.ntiles;
    if (adj) {
        // Cap the number of dp partials; each workgroup then walks several vector tiles.
        const long cap = 2L * ncu;
        if (P.ntiles > cap) P.grid = (int)cap;
        if (grid_cap > 0 && P.grid > grid_cap) P.grid = grid_cap;
        // LDS stage depth: at least two layers; the stage fits the budget.
        const size_t per_layer = (size_t)P.waves * stage_layer_bytes(P);
        const size_t budget = stage_budget(P.waves);
        const int req = P.C;  // requested layers per stage (0: automatic)
        int C = (int)(budget / per_layer);
        if (C > 16) C = 16;
        if (req >= 1 && (size_t)req * per_layer <= budget) C = req;
        const int depth = P.W <= 4 ? 4 : 2;  // layers per unrolled group of the backdiff loop
        C -= C % depth;
        if (C < depth) C = depth;
        if ((size_t)C * per_layer > budget) return MESHPROP_ERR_UNSUPPORTED;
        // Pipelined stage (conversion of a stage overlapped with the next one): two buffers of
        // an even number of layers in the same budget, and whole waves of conversion items per
        // layer (NP/2 a multiple of 64).
        P.pipe = 0;
        if ((P.W <= 4 || (P.W == 8 && P.uni)) && (P.NP / 2) % 64 == 0) {
            int Cp = (int)(budget / (2 * per_layer));
            if (Cp > 16) Cp = 16;
            if (req >= 2 && 2 * (size_t)req * per_layer <= budget) Cp = req;
            Cp -= Cp % 2;
            if (Cp >= 2) {
                P.pipe = 1;
                C = Cp;
            }
        }
        if (nn16) {
            P.pipe = 1;
            C = 1;
        }
        if (P.waves > 4 && !P.pipe) return MESHPROP_ERR_UNSUPPORTED;  // only the pipelined stage sums 8 waves
        P.C = C;
        P.sorder = (nn_adjoint(P) && nn_stage_order(P.W)) ? P.W / 2 : 0;
    } else {
        P.C = 0;
        P.pipe = 0;
    }
    size_t off = 0;
    P.off_kind = off;
    off = align256(off + (size_t)L * sizeof(int));
    P.off_a = off;
    off = align256(off + (size_t)L * P.NP * P.rowbytes);
    P.off_b = off;
    // Second table: fwddiff tangent rows, the adjoint rows of a lossy backdiff, or the
    // inverse rows of the fused forward + adjoint (whose first table holds the forward rows).
    const bool need_b = (op == MESHPROP_OP_FWDDIFF) || (op == MESHPROP_OP_BACKDIFF && !P.uni) ||
                        (op == MESHPROP_OP_FWDGRAD);
    if (need_b) off = align256(off + (size_t)L * P.NP * P.rowbytes);
    P.off_dtdp = off;
    P.off_partial = off;
    if (adj) {
        // The conversion table (dT/dp size; the K-form table takes half of it and a
        // scaled-unitary plan keeps its per-waveguide |det T| table [L][NP] behind it).
        const size_t elem = P.real ? sizeof(float) : sizeof(float2);
        off = align256(off + (size_t)L * (P.NP / 2) * P.np * 4 * elem);
        P.off_partial = off;
        off = align256(off + (size_t)P.grid * L * (P.NP / 2) * P.np * sizeof(float));
    }
    P.total = off;
    return MESHPROP_OK;
}

int check_launch() {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        g_last_hip_error = (int)e;
        return MESHPROP_ERR_HIP;
    }
    return MESHPROP_OK;
}

// Offset of the |det T| table of a scaled-unitary plan inside the conversion table region.
size_t scale_offset(const Plan& P) { return align256((size_t)P.L * P.np * (P.NP / 2) * sizeof(float4)); }

int run_impl(int op, int X, int N, int L, int B, int fft, const int* g0, const int* g1, const float* p, float* dp,
             int ldp, const float* s, int lds, const void* u_a, const void* du_a, void* u_b, void* du_b, int ldu,
             int mode, const void* target, void* ckpt, int ck_every, const meshprop_tile* tile, int flags, void* ws,
             size_t ws_bytes, hipStream_t st, int grid_cap = 0) {
    if (mode != 0) return MESHPROP_ERR_UNSUPPORTED;  // gpu/crossing.cu:122-125: only mode 0 exists
    Plan P;
    int rc = make_plan(op, X, N, L, B, fft, tile, P, flags, grid_cap);
    if (rc) return rc;
    if (B == 0) return MESHPROP_OK;
    if (ldu < N) return MESHPROP_ERR_ARG;
    if (!u_a || !u_b) return MESHPROP_ERR_ARG;
    if (L > 0 && (!g0 || (!fft && !g1) || !p)) return MESHPROP_ERR_ARG;
    if (ldp < (N / 2) * P.np) return MESHPROP_ERR_ARG;
    if (s && P.ns && lds < (N / 2) * P.ns) return MESHPROP_ERR_ARG;
    if (op == MESHPROP_OP_FWDDIFF && L > 0 && !dp) return MESHPROP_ERR_ARG;
    if (op == MESHPROP_OP_BACKDIFF && !dp) return MESHPROP_ERR_ARG;
    if (op == MESHPROP_OP_BACKDIFF && !du_a && !target) return MESHPROP_ERR_ARG;
    if (op == MESHPROP_OP_FWDGRAD && (!dp || !target)) return MESHPROP_ERR_ARG;
    if (ckpt && op == MESHPROP_OP_FWDGRAD) return MESHPROP_ERR_UNSUPPORTED;
    if ((flags & (MESHPROP_ONLY_FORWARD | MESHPROP_SKIP_FORWARD)) && op != MESHPROP_OP_FWDGRAD) return MESHPROP_ERR_ARG;
    if (ckpt && ck_every < 1) return MESHPROP_ERR_ARG;
    if (ckpt && op == MESHPROP_OP_FWDDIFF) return MESHPROP_ERR_UNSUPPORTED;
    if (ckpt && op == MESHPROP_OP_BACKDIFF && !du_b) return MESHPROP_ERR_ARG;  // carries g between segments
    if (!ckpt || ck_every > L) ck_every = L > 0 ? L : 1;
    if (P.total > 0 && (!ws || ws_bytes < P.total)) return MESHPROP_ERR_WORKSPACE;
    if (((uintptr_t)ws) & 255) return MESHPROP_ERR_WORKSPACE;
    clear_pending_hip_error();

    unsigned char* base = (unsigned char*)ws;
    int* kind = (int*)(base + P.off_kind);
    void* tab_a = base + P.off_a;
    void* tab_b = base + P.off_b;
    void* dtdp = base + P.off_dtdp;
    float* partial = (float*)(base + P.off_partial);
    float* scl = P.scaled ? (float*)(base + P.off_dtdp + scale_offset(P)) : nullptr;
    P.scales = scl;
    const float* sp = (P.ns > 0) ? s : nullptr;

    const bool do_pro = !(flags & (MESHPROP_SKIP_PROLOGUE | MESHPROP_ONLY_REDUCE | MESHPROP_ONLY_FORWARD));
    const bool do_main = !(flags & (MESHPROP_ONLY_PROLOGUE | MESHPROP_ONLY_REDUCE));
    const bool do_red = !(flags & (MESHPROP_ONLY_PROLOGUE | MESHPROP_SKIP_REDUCE | MESHPROP_ONLY_FORWARD));
    if (L > 0 && do_pro) {
        // One launch: row tables and, for the adjoint ops, the conversion table (K-form for
        // unitary crossings, dT/dp otherwise).
        int tabs = 0;
        if (op == MESHPROP_OP_FWDPROP) tabs = TAB_FWD;
        if (op == MESHPROP_OP_FWDDIFF) tabs = TAB_FWD | TAB_DIFF;
        if (op == MESHPROP_OP_BACKDIFF) tabs = TAB_INV | (P.uni ? TAB_KUNI : (TAB_ADJ | TAB_DTDP));
        if (op == MESHPROP_OP_FWDGRAD) tabs = TAB_FWD | TAB_INV_B | TAB_KUNI;
        if (P.scaled) tabs |= TAB_SCALE;
        dim3 grid((P.NP / 2 + 255) / 256, L);  // one thread per slot
        if (P.real)
            hipLaunchKernelGGL((k_rows<float2>), grid, dim3(256), 0, st, X, N, L, P.NP, P.LPV, fft, g0, g1, p, ldp, sp, lds,
                               (const float*)dp, tabs, kind, (float2*)tab_a, (float2*)tab_b, (float*)dtdp, scl, P.sorder);
        else
            hipLaunchKernelGGL((k_rows<float4>), grid, dim3(256), 0, st, X, N, L, P.NP, P.LPV, fft, g0, g1, p, ldp, sp, lds,
                               (const float*)dp, tabs, kind, (float4*)tab_a, (float4*)tab_b, (float2*)dtdp, scl, P.sorder);
        if ((rc = check_launch())) return rc;
    }

    const void* rows_b = P.uni ? tab_a : tab_b;
    const int K = ck_every;
    const int nck = L > 0 ? (L + K - 1) / K : 0;
    const size_t fstride = (size_t)B * ldu * (P.real ? sizeof(float) : 2 * sizeof(float));
    unsigned char* ck = (unsigned char*)ckpt;
    if (do_main) {
        if (op == MESHPROP_OP_FWDPROP) {
            if (!ckpt) {
                rc = launch_fwd_any(P, kind, tab_a, 0, L, u_a, u_b, ldu, st);
            } else {
                // Checkpointed forward, one launch: ckpt[c] receives the field entering layer c*K.
                rc = launch_fwd_any(P, kind, tab_a, 0, L, u_a, u_b, ldu, st, Ckpt{ckpt, K, (size_t)B * ldu});
            }
        } else if (op == MESHPROP_OP_FWDGRAD) {
            // Forward at the forward's own tile (same W, hence the same table layout), then
            // the adjoint of its output with the inverse rows of the second table.  Two
            // launches: on MI355X they ran faster than one kernel doing both per tile.
            Plan PF;
            const meshprop_tile tf = {P.W, 0, 0, 0};
            rc = make_plan(MESHPROP_OP_FWDPROP, X, N, L, B, fft, &tf, PF, flags & MESHPROP_GENERIC_KERNELS);
            if (!rc && (PF.W != P.W || PF.NP != P.NP)) rc = MESHPROP_ERR_UNSUPPORTED;
            if (!rc && !(flags & MESHPROP_SKIP_FORWARD)) {
                rc = launch_fwd_any(PF, kind, tab_a, 0, L, u_a, u_b, ldu, st);
                if (!rc) rc = check_launch();
            }
            if (!rc && !(flags & MESHPROP_ONLY_FORWARD)) {
                const size_t ps = (size_t)L * (P.NP / 2) * P.np;
                rc = launch_bwd_any(P, kind, tab_b, tab_b, dtdp, partial, ps, 0, L, u_b, nullptr, target, nullptr,
                                    du_b, ldu, st);
            }
        } else if (op == MESHPROP_OP_FWDDIFF) {
            rc = launch_fwddiff_any(P, kind, tab_a, tab_b, u_a, du_a, u_b, du_b, ldu, st);
        } else {
            const size_t ps = (size_t)L * (P.NP / 2) * P.np;
            if (!ckpt) {
                rc = launch_bwd_any(P, kind, tab_a, rows_b, dtdp, partial, ps, 0, L, u_a, du_a, target, u_b, du_b,
                                    ldu, st);
            } else if (L % P.C == 0 && K % P.C == 0) {
                // Checkpointed adjoint, one launch: every K layers the reconstructed field is
                // replaced by the exact forward field ckpt[c] (checkpoints fall on LDS stage starts).
                rc = launch_bwd_any(P, kind, tab_a, rows_b, dtdp, partial, ps, 0, L, u_a, du_a, target, u_b, du_b,
                                    ldu, st, Ckpt{ckpt, K, (size_t)B * ldu});
            } else {
                // Checkpointed adjoint, one launch per segment: segment c starts from the exact
                // forward field ckpt[c+1] (u_out for the last one); the adjoint field is carried in
                // dJdu_in.
                for (int c = nck - 1; c >= 0 && !rc; --c) {
                    const int l0 = c * K, nl = (L - l0 < K) ? L - l0 : K;
                    const bool last = (c == nck - 1);
                    const void* us = last ? u_a : (const void*)(ck + (size_t)(c + 1) * fstride);
                    rc = launch_bwd_any(P, kind, tab_a, rows_b, dtdp, partial, ps, l0, nl, us, last ? du_a : du_b,
                                        last ? target : nullptr, c == 0 ? u_b : nullptr, du_b, ldu, st);
                }
            }
        }
        if (rc) return rc;
        if ((rc = check_launch())) return rc;
    }
    if ((op == MESHPROP_OP_BACKDIFF || op == MESHPROP_OP_FWDGRAD) && L > 0 && do_red) {
        const long total = (long)L * (P.sorder > 0 ? P.NP / 2 : N / 2) * P.np;
        const size_t E = (size_t)L * (P.NP / 2) * P.np;
        hipLaunchKernelGGL(k_dp_reduce, dim3((unsigned)((total + 63) / 64)), dim3(256), 0, st, L, N / 2, P.NP / 2,
                           P.np, P.grid, E, (const float*)partial, dp, ldp, (flags & MESHPROP_DP_ASSIGN) ? 1 : 0,
                           P.sorder);
        if ((rc = check_launch())) return rc;
    }
    return MESHPROP_OK;
}

// ---- static scratch of meshprop_run_noalloc --------------------------------------------
// Its device address differs per device; host code never uses the array's name as an
// address (it is a host-side stand-in), only the pointer hipGetSymbolAddress returns.
constexpr size_t kScratchBytes = MESHPROP_NOALLOC_SCRATCH_BYTES;
static_assert(kScratchBytes % 256 == 0, "scratch size keeps the workspace alignment");
constexpr int kMaxDevices = 64;

std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

// The scratch's address on the current device, looked up once per device (with that device
// current) on its first call.
int scratch_on_current_device(void** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        g_last_hip_error = (int)e;
        return MESHPROP_ERR_HIP;
    }
    if (dev < 0 || dev >= kMaxDevices) return MESHPROP_ERR_UNSUPPORTED;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[dev] == nullptr) {
        void* ptr = nullptr;
        e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || ptr == nullptr) {
            g_last_hip_error = (int)e;
            return MESHPROP_ERR_HIP;
        }
        if (((uintptr_t)ptr) & 255) return MESHPROP_ERR_WORKSPACE;
        g_scratch_ptr[dev] = ptr;
    }
    *out = g_scratch_ptr[dev];
    return MESHPROP_OK;
}

// Workgroup cap under which a call's workspace fits the scratch: 0 when it fits as planned,
// -1 when it cannot fit (row tables alone too large, or no partial slab fits).
int scratch_grid_cap(int op, int X, int N, int L, int B, int fft, const meshprop_tile* tile, int flags, int& rc) {
    Plan P;
    rc = make_plan(op, X, N, L, B, fft, tile, P, flags);
    if (rc) return -1;
    if (P.total <= kScratchBytes) return 0;
    const bool adj = (op == MESHPROP_OP_BACKDIFF || op == MESHPROP_OP_FWDGRAD);
    if (!adj || P.off_partial >= kScratchBytes) return -1;
    const size_t slab = (size_t)L * (P.NP / 2) * P.np * sizeof(float);  // one workgroup's partial
    const size_t cap = (kScratchBytes - P.off_partial) / slab;  // a multiple of 256 bytes stays in range
    return cap >= 1 ? (int)std::min<size_t>(cap, (size_t)P.grid) : -1;
}

}  // namespace

extern "C" {

int meshprop_run_noalloc(int op, int X, int N, int L, int B, int fft, const int* geom0, const int* geom1,
                         const float* p, float* dp, int ldp, const float* s, int lds, const void* u_a,
                         const void* du_a, void* u_b, void* du_b, int ldu, int mode, const void* target, void* ckpt,
                         int ckpt_every, const meshprop_tile* tile, int flags, void* stream) {
    int rc = MESHPROP_OK;
    const int cap = scratch_grid_cap(op, X, N, L, B, fft, tile, flags, rc);
    if (rc) return rc;
    if (cap < 0) return MESHPROP_ERR_SCRATCH;
    void* ws = nullptr;
    if ((rc = scratch_on_current_device(&ws))) return rc;
    return run_impl(op, X, N, L, B, fft, geom0, geom1, p, dp, ldp, s, lds, u_a, du_a, u_b, du_b, ldu, mode, target,
                    ckpt, ckpt_every, tile, flags, ws, kScratchBytes, (hipStream_t)stream, cap);
}

size_t meshprop_workspace_size(int op, int X, int N, int L, int B, int fft, const meshprop_tile* tile) {
    Plan P;
    size_t total = 0;
    // Enough for every plan the flags of a call can select: the scaled-unitary plan of generic
    // crossings (MESHPROP_SCALED_UNITARY) and the general kernels (MESHPROP_GENERIC_KERNELS, whose
    // workgroups can be smaller, hence more dp partials).
    for (const int fl : {0, (int)MESHPROP_GENERIC_KERNELS, (int)MESHPROP_SCALED_UNITARY,
                         (int)(MESHPROP_SCALED_UNITARY | MESHPROP_GENERIC_KERNELS)}) {
        if ((fl & MESHPROP_SCALED_UNITARY) && X != X_GENERIC) continue;
        if (make_plan(op, X, N, L, B, fft, tile, P, fl) == MESHPROP_OK) total = std::max(total, P.total);
    }
    return total;
}

int meshprop_resolve_tile(int op, int X, int N, int B, meshprop_tile* out) {
    Plan P;
    int rc = make_plan(op, X, N, 1, B, 0, out, P);
    if (rc) return rc;
    out->w = P.W;
    out->v = P.V;
    out->waves = P.waves;
    out->chunk = P.C;
    return MESHPROP_OK;
}

int meshprop_adjoint_kernel(int op, int X, int N, int L, int B, int fft, const meshprop_tile* tile, int flags) {
    Plan P;
    const int rc = make_plan(op, X, N, L, B, fft, tile, P, flags);
    if (rc) return -rc;
    if (op != MESHPROP_OP_BACKDIFF && op != MESHPROP_OP_FWDGRAD) return 0;
    return nn_adjoint(P) ? 1 : 0;
}

int meshprop_forward_kernel(int op, int X, int N, int L, int B, int fft, const meshprop_tile* tile, int flags) {
    Plan P;
    int rc = make_plan(op, X, N, L, B, fft, tile, P, flags);
    if (rc) return -rc;
    if (op == MESHPROP_OP_FWDGRAD) {
        // the forward of fwd+grad runs at the adjoint's width (as run_impl plans it)
        const meshprop_tile tf = {P.W, 0, 0, 0};
        Plan PF;
        rc = make_plan(MESHPROP_OP_FWDPROP, X, N, L, B, fft, &tf, PF, flags & MESHPROP_GENERIC_KERNELS);
        if (rc) return -rc;
        return nn_forward(PF, PF.ncu) ? 1 : 0;
    }
    if (op == MESHPROP_OP_FWDDIFF) return nn_fwddiff(P) ? 1 : 0;
    if (op != MESHPROP_OP_FWDPROP) return 0;
    return nn_forward(P, P.ncu) ? 1 : 0;
}

int meshprop_run(int op, int X, int N, int L, int B, int fft, const int* geom0, const int* geom1, const float* p,
                 float* dp, int ldp, const float* s, int lds, const void* u_a, const void* du_a, void* u_b, void* du_b,
                 int ldu, int mode, const void* target, void* ckpt, int ckpt_every, const meshprop_tile* tile,
                 int flags, void* workspace, size_t workspace_bytes, void* stream) {
    return run_impl(op, X, N, L, B, fft, geom0, geom1, p, dp, ldp, s, lds, u_a, du_a, u_b, du_b, ldu, mode, target,
                    ckpt, ckpt_every, tile, flags, workspace, workspace_bytes, (hipStream_t)stream);
}

int meshprop_fwdprop(int X, int N, int L, int B, const int* lens, const int* shifts, const float* p, int ldp,
                     const float* s, int lds, const void* u_in, void* u_out, int ldu, int mode, void* workspace,
                     size_t workspace_bytes, void* stream) {
    return run_impl(MESHPROP_OP_FWDPROP, X, N, L, B, 0, lens, shifts, p, nullptr, ldp, s, lds, u_in, nullptr, u_out,
                    nullptr, ldu, mode, nullptr, nullptr, 0, nullptr, 0, workspace, workspace_bytes,
                    (hipStream_t)stream);
}

int meshprop_fwdprop_ft(int X, int N, int L, int B, const int* strides, const float* p, int ldp, const float* s,
                        int lds, const void* u_in, void* u_out, int ldu, int mode, void* workspace,
                        size_t workspace_bytes, void* stream) {
    return run_impl(MESHPROP_OP_FWDPROP, X, N, L, B, 1, strides, nullptr, p, nullptr, ldp, s, lds, u_in, nullptr,
                    u_out, nullptr, ldu, mode, nullptr, nullptr, 0, nullptr, 0, workspace, workspace_bytes,
                    (hipStream_t)stream);
}

int meshprop_fwddiff(int X, int N, int L, int B, const int* lens, const int* shifts, const float* p, const float* dp,
                     int ldp, const float* s, int lds, const void* u_in, const void* du_in, void* u_out, void* du_out,
                     int ldu, int mode, void* workspace, size_t workspace_bytes, void* stream) {
    return run_impl(MESHPROP_OP_FWDDIFF, X, N, L, B, 0, lens, shifts, p, (float*)dp, ldp, s, lds, u_in, du_in, u_out,
                    du_out, ldu, mode, nullptr, nullptr, 0, nullptr, 0, workspace, workspace_bytes, (hipStream_t)stream);
}

int meshprop_fwddiff_ft(int X, int N, int L, int B, const int* strides, const float* p, const float* dp, int ldp,
                        const float* s, int lds, const void* u_in, const void* du_in, void* u_out, void* du_out,
                        int ldu, int mode, void* workspace, size_t workspace_bytes, void* stream) {
    return run_impl(MESHPROP_OP_FWDDIFF, X, N, L, B, 1, strides, nullptr, p, (float*)dp, ldp, s, lds, u_in, du_in,
                    u_out, du_out, ldu, mode, nullptr, nullptr, 0, nullptr, 0, workspace, workspace_bytes,
                    (hipStream_t)stream);
}

int meshprop_backdiff(int X, int N, int L, int B, const int* lens, const int* shifts, const float* p, float* dp,
                      int ldp, const float* s, int lds, const void* u_out, const void* dJdu_out, void* u_in,
                      void* dJdu_in, int ldu, int mode, void* workspace, size_t workspace_bytes, void* stream) {
    return run_impl(MESHPROP_OP_BACKDIFF, X, N, L, B, 0, lens, shifts, p, dp, ldp, s, lds, u_out, dJdu_out, u_in,
                    dJdu_in, ldu, mode, nullptr, nullptr, 0, nullptr, 0, workspace, workspace_bytes,
                    (hipStream_t)stream);
}

int meshprop_backdiff_ft(int X, int N, int L, int B, const int* strides, const float* p, float* dp, int ldp,
                         const float* s, int lds, const void* u_out, const void* dJdu_out, void* u_in, void* dJdu_in,
                         int ldu, int mode, void* workspace, size_t workspace_bytes, void* stream) {
    return run_impl(MESHPROP_OP_BACKDIFF, X, N, L, B, 1, strides, nullptr, p, dp, ldp, s, lds, u_out, dJdu_out, u_in,
                    dJdu_in, ldu, mode, nullptr, nullptr, 0, nullptr, 0, workspace, workspace_bytes,
                    (hipStream_t)stream);
}

int meshprop_last_hip_error(void) { return g_last_hip_error; }

const char* meshprop_status_string(int status) {
    switch (status) {
        case MESHPROP_OK: return "ok";
        case MESHPROP_ERR_ARG: return "invalid argument";
        case MESHPROP_ERR_WORKSPACE: return "workspace missing, too small or not 256-byte aligned";
        case MESHPROP_ERR_UNSUPPORTED: return "unsupported configuration";
        case MESHPROP_ERR_HIP: return "HIP launch error";
        case MESHPROP_ERR_SCRATCH: return "call needs more device memory than the static scratch of meshprop_run_noalloc";
        default: return "unknown status";
    }
}

const char* meshprop_version(void) { return "meshprop-gfx950 0.1"; }

}  // extern "C"
