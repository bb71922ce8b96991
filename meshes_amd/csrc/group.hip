// group.hip -- grouped launch: many meshes of different width, depth and geometry propagated
// by one persistent kernel that takes (mesh, partial slot) work items from a device queue.
//
// The plan lives at the start of the workspace (header, per-mesh descriptors, item lists);
// the prologue, the persistent kernel and the dp reduction read it from there and walk their
// work grid-stride, so their launch shapes depend on the device alone and one captured
// hipGraph replays every plan uploaded into the same workspace.
#include <algorithm>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "kernels.h"

namespace mp {

// Kernel families, each a persistent kernel of four-wave workgroups: family 0 holds the tiles
// with W <= 8 (two workgroups per CU); family 1 the W = 16 tile of NN meshes (LDS-staged
// forward, NN-specialised adjoint; compiled for two waves per SIMD, two workgroups per CU);
// family 2 the W = 16 tile of the other meshes (general forward and adjoint, whose registers
// allow one wave per SIMD).  Kept apart, the general W = 16 code does not hold family 1 to one
// wave per SIMD.  Variant v of the persistent kernels: (W, V) = kGroupTiles[v].
struct GTile {
    int w, v, fam;
};
constexpr GTile kGroupTiles[] = {{2, 1, 0}, {2, 2, 0}, {2, 4, 0}, {4, 1, 0}, {4, 2, 0}, {8, 1, 0}, {16, 1, 1}};
constexpr int kGroupVariants = sizeof(kGroupTiles) / sizeof(kGroupTiles[0]);
constexpr int kGroupFams = 3;
constexpr int kFamWaves[kGroupFams] = {4, 4, 4};
static_assert(kFamWaves[0] <= kStageWaves && kFamWaves[2] <= kStageWaves, "stage conversion sums at most kStageWaves waves");
constexpr size_t kGroupStage = 64 * 1024;  // LDS of one workgroup (backdiff stage)
constexpr size_t kGroupLds = kGroupStage + 16;  // stage + the queue slot

enum { G_FWD = 1, G_GRAD = 2 };  // what an item runs
enum { B_KIND = 1, B_FWD = 2, B_INV = 4, B_DTDP = 8 };

// Row tables, K tables and fields are typed for complex meshes (float4 rows, float2 fields);
// the descriptor of a real (orth) mesh holds the same buffers as float2 rows, float K
// coefficients and float fields, read through casts where real is set.
struct GMesh {
    int what, var, X, N, L, B, fft, LPV, NP, C, np, ns, ldp, lds, ldu, S, build, assign;
    int real;
    int nnadj;  // gradient items run the NN-specialised adjoint (NN meshes, one 64-lane vector group)
    int kso;    // > 0: K table and dp partials in the NN adjoint's stage order (W >= 4 NN meshes)
    long ntiles, E;
    long pro_chunks;  // prologue blocks (of 256 slots) per layer
    const int *g0, *g1;
    const float *p, *s;
    float* dp;
    int* kind;
    float4 *rows_fwd, *rows_inv;
    float2* dtdp;
    float* partial;
    const float2 *u_in, *target, *g_out;
    float2 *u_out, *g_in, *u_rec;
};

struct GHeader {
    int nmesh, pad0;
    int nitems[kGroupFams];
    int grab[kGroupFams], done[kGroupFams];  // queue counters; the last workgroup out of a launch resets them
    long pro_total, red_total;
    long off_pro, off_red, off_mesh, off_items[kGroupFams];
};

// Descriptors are read through the constant address space: they do not change while a launch
// runs, their loads become scalar, and pointers loaded from constant memory are known to point
// to global memory (generic pointers would compile to flat instructions, which also count
// against the LDS wait counter and serialise the row prefetch with every LDS access).
typedef __attribute__((address_space(4))) const GMesh KMesh;
typedef __attribute__((address_space(4))) const long KLong;
typedef __attribute__((address_space(4))) const int KInt;  // layer kinds: written by the prologue only

__device__ __forceinline__ int find_mesh(KLong* begin, const int nmesh, const long u) {
    // Largest e with begin[e] <= u (entries without units have begin[e] == begin[e + 1]).
    int lo = 0, hi = nmesh - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (begin[mid] <= u)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

__global__ void __launch_bounds__(256) k_group_prologue(const unsigned char* __restrict__ ws) {
    const GHeader* H = reinterpret_cast<const GHeader*>(ws);
    KLong* begin = (KLong*)(ws + H->off_pro);
    KMesh* meshes = (KMesh*)(ws + H->off_mesh);
    for (long u = blockIdx.x; u < H->pro_total; u += gridDim.x) {
        const int e = find_mesh(begin, H->nmesh, u);
        KMesh& M = meshes[e];
        const long local = u - begin[e];
        const int l = (int)(local / M.pro_chunks);
        const int m = (int)(local % M.pro_chunks) * 256 + threadIdx.x;
        if (m >= M.NP / 2) continue;
        // One crossing evaluation per slot for all of the entry's tables.
        int tabs = 0;
        float4* ta = nullptr;
        float4* tb = nullptr;
        if (M.build & B_FWD) {
            tabs |= TAB_FWD;
            ta = M.rows_fwd;
            if (M.build & B_INV) {
                tabs |= TAB_INV_B;
                tb = M.rows_inv;
            }
        } else if (M.build & B_INV) {
            tabs |= TAB_INV;
            ta = M.rows_inv;
        }
        if (M.build & B_DTDP) tabs |= TAB_KUNI;  // grouped meshes are unitary: K-form
        if (M.real)
            slot_tables_at<float2>(M.X, M.N, M.NP, M.LPV, M.fft, M.g0, M.g1, M.p, M.ldp, M.s, M.lds, nullptr, tabs,
                                   M.kind, reinterpret_cast<float2*>(ta), reinterpret_cast<float2*>(tb),
                                   reinterpret_cast<float*>(M.dtdp), nullptr, l, m, M.kso);
        else
            slot_tables_at<float4>(M.X, M.N, M.NP, M.LPV, M.fft, M.g0, M.g1, M.p, M.ldp, M.s, M.lds, nullptr, tabs,
                                   M.kind, ta, tb, M.dtdp, nullptr, l, m, M.kso);
    }
}

// dp[l*ldp + m*np + k] += sum over slots s of partial[s*E + (l*NPh + m)*np + k], m < N/2,
// summed in slot order (four interleaved partial sums), so the result is deterministic.
__global__ void __launch_bounds__(256) k_group_reduce(const unsigned char* __restrict__ ws) {
    // Unit u of a mesh = 64 consecutive dp entries; wave q of the block sums the slots
    // [q*S/4, (q+1)*S/4) with four interleaved accumulators, and the four wave sums are added
    // in wave order through LDS (a fixed order: bitwise reproducible).
    __shared__ float red[4][64];
    const GHeader* H = reinterpret_cast<const GHeader*>(ws);
    KLong* begin = (KLong*)(ws + H->off_red);
    KMesh* meshes = (KMesh*)(ws + H->off_mesh);
    const int lane = threadIdx.x & 63, wq = threadIdx.x >> 6;
    for (long u = blockIdx.x; u < H->red_total; u += gridDim.x) {
        const int e = find_mesh(begin, H->nmesh, u);
        KMesh& M = meshes[e];
        const long idx = (u - begin[e]) * 64 + lane;
        const int Nh = M.N / 2;
        // Stage-order partials (kso > 0): units run over positions (coalesced reads) and
        // position p*64 + t lands in slot t*kso + p.
        const int Nr = M.kso > 0 ? M.NP / 2 : Nh;
        bool ok = idx < (long)M.L * Nr * M.np;
        int k = 0, m = 0, l = 0, pos = 0;
        if (ok) {
            k = (int)(idx % M.np);
            const long lm = idx / M.np;
            pos = (int)(lm % Nr);
            l = (int)(lm / Nr);
            m = stage_slot(pos, M.kso);
            ok = m < Nh;
        }
        float acc = 0.f;
        if (ok) {
            const float* src = M.partial + ((size_t)l * (M.NP / 2) + pos) * M.np + k;
            const size_t E = (size_t)M.E;
            const int s0 = (int)((long)M.S * wq / 4), s1 = (int)((long)M.S * (wq + 1) / 4);
            // Eight interleaved accumulators: eight loads in flight per lane (the sums are
            // latency-bound otherwise: a large mesh has hundreds of slots).
            float a[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
            int s = s0;
            for (; s + 7 < s1; s += 8)
#pragma unroll
                for (int r = 0; r < 8; ++r) a[r] += src[(size_t)(s + r) * E];
            for (; s < s1; ++s) a[0] += src[(size_t)s * E];
            acc = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
        }
        red[wq][lane] = acc;
        __syncthreads();
        if (wq == 0 && ok) {
            const float sum = ((red[0][lane] + red[1][lane]) + red[2][lane]) + red[3][lane];
            float* d = M.dp + (size_t)l * M.ldp + (size_t)m * M.np + k;
            *d = M.assign ? sum : *d + sum;
        }
        __syncthreads();  // red is rewritten by the next unit
    }
}

// One work item: the vector tiles slot, slot + S, ... of one mesh, each through the whole
// depth: forward, then (gradient items) the adjoint with inverse reconstruction; the adjoint
// of a tile reads back the output its own threads stored.
// The forward tile runs as a called function: inlined beside the adjoint it raises the
// kernel's register allocation above 256 (one wave per SIMD instead of two).  Its arguments
// arrive in vector registers and are made scalar again with readfirstlane.
// W = 16 NN meshes (kso > 0: one 64-lane vector per wave) run the forward with their rows
// staged through LDS (fwd_tile_nn, its three-slot ring in the stage buffer, which no adjoint
// uses until the forward's last barrier): each wave reading 16 KiB of rows per layer through
// the address unit bounds the general tile.
// T: float2 (complex fields) or float (real fields, orth).
template <class T, int W, int V, bool NN>
__device__ __noinline__ void group_fwd_tile(KMesh* Mp, const long tile_v, float4* stage) {
    typedef typename RowOf<T>::type R;
    KMesh& M = *uniform_ptr(Mp);
    const long tile = (long)__builtin_amdgcn_readfirstlane((unsigned)tile_v);
    const R* rows = reinterpret_cast<const R*>(M.rows_fwd);
    const T* u_in = reinterpret_cast<const T*>(M.u_in);
    T* u_out = reinterpret_cast<T*>(M.u_out);
    if constexpr (NN) {
        static_assert(W == 16 && V == 1, "the NN family holds the W = 16 tile");
        static_assert(3 * 64 * W * sizeof(float4) <= kGroupStage, "forward row ring inside the stage buffer");
        static_assert(kFamWaves[1] == 4, "fwd_tile_nn copies rows with 256 threads");
        fwd_tile_nn<T, 16, 1>(M.N, 0, M.L, M.B, (KInt*)M.kind, rows, u_in, u_out, M.ldu, tile, Ckpt{nullptr, 0, 0},
                              reinterpret_cast<R*>(stage));
    } else {
        fwd_tile<T, W, V, KInt*>(M.N, 0, M.L, M.B, M.LPV, M.NP, (KInt*)M.kind, rows, u_in, u_out, M.ldu, tile);
    }
}

// NN: the W = 16 NN family (one-layer stages in stage order).  In family 0, NN meshes of one
// 64-lane vector group (nnadj) run the NN-specialised adjoint as the plain path does: stages
// of 16 / W layers, four waves (64 KiB of LDS), entries in stage order at W >= 4.
template <class T, int W, int V, bool NN = false>
__device__ __forceinline__ void group_item(KMesh& M, const int slot, float4* stage_raw) {
    typedef typename RowOf<T>::type R;
    float* part = (M.what & G_GRAD) ? M.partial + (size_t)slot * M.E : nullptr;
    const R* rows_inv = reinterpret_cast<const R*>(M.rows_inv);
    const T* ktab = reinterpret_cast<const T*>(M.dtdp);
    const T* u_out = reinterpret_cast<const T*>(M.u_out);
    const T* g_out = reinterpret_cast<const T*>(M.g_out);
    const T* target = reinterpret_cast<const T*>(M.target);
    T* u_rec = reinterpret_cast<T*>(M.u_rec);
    T* g_in = reinterpret_cast<T*>(M.g_in);
    R* stage = reinterpret_cast<R*>(stage_raw);
    bool first = true;
    for (long tile = slot; tile < M.ntiles; tile += M.S) {
        if (M.what & G_FWD)
            group_fwd_tile<T, W, V, NN>(&M, tile, stage_raw);
        if (M.what & G_GRAD) {
            if constexpr (NN) {
                bwd_tile_nn<T, 16, 1, kFamWaves[1], 1, false>(M.N, 0, M.L, M.B, (KInt*)M.kind, rows_inv, ktab, part,
                                                              u_out, g_out, target, u_rec, g_in, M.ldu, tile, first,
                                                              stage, Ckpt{nullptr, 0, 0}, nullptr);
            } else {
                if constexpr (W <= 8) {
                    static_assert(nn_lds_bytes(W, kFamWaves[0], nn_stage_depth(W), false, sizeof(float4)) <=
                                      kGroupStage,
                                  "NN adjoint stage inside the group stage");
                    if (M.nnadj) {
                        bwd_tile_nn<T, W, V, kFamWaves[0], nn_stage_depth(W), false, false>(
                            M.N, 0, M.L, M.B, (KInt*)M.kind, rows_inv, ktab, part, u_out, g_out, target, u_rec, g_in,
                            M.ldu, tile, first, stage, Ckpt{nullptr, 0, 0}, nullptr);
                        first = false;
                        continue;
                    }
                }
                bwd_tile<T, W, V, true, KInt*>(M.N, 0, M.L, M.B, M.LPV, M.NP, M.C, (KInt*)M.kind, rows_inv, rows_inv,
                                               ktab, part, u_out, g_out, target, u_rec, g_in, M.ldu, tile, first,
                                               stage);
            }
        }
        first = false;
    }
}

template <int FAM>
__global__ void __launch_bounds__(64 * kFamWaves[FAM]) __attribute__((amdgpu_waves_per_eu(FAM <= 1 ? 2 : 1)))
k_group(unsigned char* __restrict__ ws) {
    GHeader* H = reinterpret_cast<GHeader*>(ws);
    KMesh* meshes = (KMesh*)(ws + H->off_mesh);
    KInt* items = (KInt*)(ws + H->off_items[FAM]);  // (mesh, slot) pairs
    const int n = H->nitems[FAM];
    extern __shared__ unsigned char smem_raw[];
    float4* stage = reinterpret_cast<float4*>(smem_raw);
    // The queue slot lives behind the stage in the dynamic block: a static __shared__ variable
    // would move the stage off LDS address 0, which made the adjoint about 1.4x slower.
    int& s_next = *reinterpret_cast<int*>(smem_raw + kGroupStage);
    // The first item of workgroup b is item b: the dispatcher spreads consecutive workgroups
    // over distinct CUs, so the costliest items (the list is sorted by cost) start one per CU
    // instead of wherever the first atomics land.  Later items come from the queue.
    int it = blockIdx.x;
    for (;;) {
        if (it >= n) break;
        const int wm = items[2 * it], ws_slot = items[2 * it + 1];
        KMesh& M = meshes[wm];
        if constexpr (FAM == 0) {
            if (M.real) {
                switch (M.var) {
                    case 0: group_item<float, 2, 1>(M, ws_slot, stage); break;
                    case 1: group_item<float, 2, 2>(M, ws_slot, stage); break;
                    case 2: group_item<float, 2, 4>(M, ws_slot, stage); break;
                    case 3: group_item<float, 4, 1>(M, ws_slot, stage); break;
                    case 4: group_item<float, 4, 2>(M, ws_slot, stage); break;
                    case 5: group_item<float, 8, 1>(M, ws_slot, stage); break;
                    default: break;
                }
            } else {
                switch (M.var) {
                    case 0: group_item<float2, 2, 1>(M, ws_slot, stage); break;
                    case 1: group_item<float2, 2, 2>(M, ws_slot, stage); break;
                    case 2: group_item<float2, 2, 4>(M, ws_slot, stage); break;
                    case 3: group_item<float2, 4, 1>(M, ws_slot, stage); break;
                    case 4: group_item<float2, 4, 2>(M, ws_slot, stage); break;
                    case 5: group_item<float2, 8, 1>(M, ws_slot, stage); break;
                    default: break;
                }
            }
        } else if constexpr (FAM == 1) {
            group_item<float2, 16, 1, true>(M, ws_slot, stage);
        } else {
            if (M.real)
                group_item<float, 16, 1>(M, ws_slot, stage);
            else
                group_item<float2, 16, 1>(M, ws_slot, stage);
        }
        if (threadIdx.x == 0) s_next = (int)gridDim.x + atomicAdd(&H->grab[FAM], 1);
        __syncthreads();
        it = __builtin_amdgcn_readfirstlane(s_next);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        __threadfence();
        if (atomicAdd(&H->done[FAM], 1) == (int)gridDim.x - 1) {
            // Every workgroup has taken its last item: reset the queue for the next launch.
            atomicExch(&H->grab[FAM], 0);
            atomicExch(&H->done[FAM], 0);
        }
    }
}

}  // namespace mp

using namespace mp;

struct meshprop_group {
    struct Info {
        int W, V, LPV, NP, C, S, var, fam, what, build, np, ns, nnadj, kso, real;
        long ntiles, E, pro_units, red_units;
        bool own_tables;
        size_t off_tab, off_partial;  // workspace offsets (tables when own_tables)
        size_t tab_kind, tab_fwd, tab_inv, tab_dtdp;  // offsets inside the table block
    };
    std::vector<meshprop_group_entry> ent;
    std::vector<Info> info;
    std::vector<int2> items[kGroupFams];
    size_t header_bytes = 0, total = 0;
    long pro_total = 0, red_total = 0;
    int ncu = 256, grid[kGroupFams] = {0, 0, 0}, dev = 0;
    int sig = 0;  // MESHPROP_GROUP_SIG_* bits: which launches the plan needs
    mutable std::vector<unsigned char> image;
};

namespace {

inline size_t a256(size_t x) { return (x + 255) & ~(size_t)255; }

int group_var(int W, int V) {
    for (int i = 0; i < kGroupVariants; ++i)
        if (kGroupTiles[i].w == W && kGroupTiles[i].v == V) return i;
    return -1;
}

int group_default_w(int N) { return N <= 128 ? 2 : N <= 256 ? 4 : N <= 512 ? 8 : 16; }

int group_w(int N, const meshprop_tile& t) {
    const int w = t.w;
    if ((w == 2 || w == 4 || w == 8 || w == 16) && (long)w * 64 >= N) return w;
    return group_default_w(N);
}

int lpv_of(int N, int W) {
    int p = 1;
    while (p < (N + W - 1) / W) p <<= 1;
    return p;
}

// Table block of one entry: kind[L] | fwd rows [L][NP] | inverse rows [L][NP] | dT/dp.
void table_layout(int N, int L, int W, size_t& kind, size_t& fwd, size_t& inv, size_t& dtdp, size_t& total) {
    const int NP = lpv_of(N, W) * W;
    kind = 0;
    fwd = a256((size_t)L * sizeof(int));
    inv = a256(fwd + (size_t)L * NP * sizeof(float4));
    dtdp = a256(inv + (size_t)L * NP * sizeof(float4));
    total = a256(dtdp + (size_t)L * (NP / 2) * 2 * 4 * sizeof(float2));
}

int cu_count() {
    int dev = 0, n = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 256;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) return 256;
    return n;
}

// Per-device launch shape of the grouped kernels (resident workgroups per family) and the
// second stream the W = 16 families run on, set up once per device.
struct DeviceGroupConfig {
    std::once_flag once;
    int ncu = 256, grid[kGroupFams] = {256, 256, 256};
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    std::mutex lock;  // orders fork / join sequences of concurrent host threads
    bool ok = false;
};
constexpr int kMaxDevices = 64;
DeviceGroupConfig g_dev_cfg[kMaxDevices];

DeviceGroupConfig* device_config(int dev) {
    if (dev < 0 || dev >= kMaxDevices) return nullptr;
    DeviceGroupConfig& c = g_dev_cfg[dev];
    std::call_once(c.once, [&c]() {
        c.ncu = cu_count();
        const void* fn[kGroupFams] = {(const void*)k_group<0>, (const void*)k_group<1>, (const void*)k_group<2>};
        bool attr = true;
        for (int fam = 0; fam < kGroupFams; ++fam) {
            // The stage and queue slot exceed the default dynamic LDS limit of 64 KiB.
            attr = attr && hipFuncSetAttribute(fn[fam], hipFuncAttributeMaxDynamicSharedMemorySize, (int)kGroupLds) ==
                               hipSuccess;
            int nb = 0;
            hipError_t rc = fam == 0   ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_group<0>, 64 * kFamWaves[0],
                                                                                      kGroupLds)
                            : fam == 1 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_group<1>, 64 * kFamWaves[1],
                                                                                      kGroupLds)
                                       : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_group<2>, 64 * kFamWaves[2],
                                                                                      kGroupLds);
            if (rc != hipSuccess || nb < 1) nb = 1;
            c.grid[fam] = nb * c.ncu;
        }
        c.ok = attr && hipStreamCreateWithFlags(&c.side, hipStreamNonBlocking) == hipSuccess &&
               hipEventCreateWithFlags(&c.ev_fork, hipEventDisableTiming) == hipSuccess &&
               hipEventCreateWithFlags(&c.ev_join, hipEventDisableTiming) == hipSuccess;
    });
    return &c;
}

int check_entry(const meshprop_group_entry& e) {
    if (e.op != MESHPROP_OP_FWDPROP && e.op != MESHPROP_OP_BACKDIFF && e.op != MESHPROP_OP_FWDGRAD)
        return MESHPROP_ERR_ARG;
    if (e.X < 0 || e.X > 3) return MESHPROP_ERR_ARG;
    if (e.op != MESHPROP_OP_FWDPROP && e.X == X_GENERIC) return MESHPROP_ERR_UNSUPPORTED;
    if (e.N <= 0 || (e.N & 1) || e.L < 0 || e.B < 0) return MESHPROP_ERR_ARG;
    if (e.N > 1024) return MESHPROP_ERR_UNSUPPORTED;
    const int np = n_phase_of(e.X), ns = n_split_of(e.X);
    if (e.L > 0 && (!e.geom0 || (!e.fft && !e.geom1) || !e.p)) return MESHPROP_ERR_ARG;
    if (e.ldp < (e.N / 2) * np) return MESHPROP_ERR_ARG;
    if (e.s && ns && e.lds < (e.N / 2) * ns) return MESHPROP_ERR_ARG;
    if (e.B > 0) {
        if (e.ldu < e.N || !e.u_out) return MESHPROP_ERR_ARG;
        if (e.op != MESHPROP_OP_BACKDIFF && !e.u_in) return MESHPROP_ERR_ARG;
        if (e.op != MESHPROP_OP_FWDPROP && (!e.dp || (!e.target && !e.dJdu_out))) return MESHPROP_ERR_ARG;
    }
    if (((uintptr_t)e.tables) & 255) return MESHPROP_ERR_ARG;
    return MESHPROP_OK;
}

// Offsets of the families' item lists in the plan image (after header, prologue and reduction
// unit tables and the mesh descriptors); returns the image size.
size_t item_offsets(const meshprop_group& g, size_t (&it)[kGroupFams]) {
    const int n = (int)g.ent.size();
    const size_t hdr = a256(sizeof(GHeader));
    const size_t pro = hdr, red = a256(pro + (size_t)(n + 1) * sizeof(long));
    const size_t mesh = a256(red + (size_t)(n + 1) * sizeof(long));
    size_t off = a256(mesh + (size_t)n * sizeof(GMesh));
    for (int fam = 0; fam < kGroupFams; ++fam) {
        it[fam] = off;
        off = a256(off + g.items[fam].size() * sizeof(int2));
    }
    return off;
}

int plan_group(meshprop_group& g) {
    if (hipGetDevice(&g.dev) != hipSuccess) g.dev = 0;
    const DeviceGroupConfig* dc = device_config(g.dev);
    g.ncu = dc ? dc->ncu : 256;
    const int n = (int)g.ent.size();
    g.info.resize(n);
    struct Cand {
        double cost;
        int e, s, fam;
    };
    std::vector<Cand> cand;
    size_t off = 0;
    g.pro_total = g.red_total = 0;
    for (int i = 0; i < n; ++i) {
        const meshprop_group_entry& e = g.ent[i];
        meshprop_group::Info& f = g.info[i];
        f.W = group_w(e.N, e.tile);
        f.LPV = lpv_of(e.N, f.W);
        f.NP = f.LPV * f.W;
        f.np = n_phase_of(e.X);
        f.ns = n_split_of(e.X);
        f.what = (e.op == MESHPROP_OP_FWDPROP) ? G_FWD : (e.op == MESHPROP_OP_BACKDIFF) ? G_GRAD : (G_FWD | G_GRAD);
        const int G = 64 / f.LPV;
        f.V = 1;
        if (e.tile.v >= 1 && group_var(f.W, e.tile.v) >= 0) {
            f.V = e.tile.v;
        } else {
            for (int v = 4; v >= 1; v >>= 1) {  // largest tile that still gives two waves per SIMD
                if (group_var(f.W, v) < 0) continue;
                if (((long)e.B + (long)G * v - 1) / ((long)G * v) >= 8L * g.ncu) {
                    f.V = v;
                    break;
                }
            }
        }
        f.var = group_var(f.W, f.V);
        // NN meshes of unitary crossings with one 64-lane vector group run the NN-specialised
        // adjoint (the W = 16 ones as family 1, with one-layer stages): K table and partials in
        // its stage order at W >= 4.
        f.nnadj = (!e.fft && f.LPV == 64 && (e.X == X_MZI || e.X == X_SYM) && (f.W < 16 || f.V == 1)) ? 1 : 0;
        f.real = (e.X == X_ORTH) ? 1 : 0;  // real meshes run the general adjoint (families 0 and 2)
        f.kso = (f.nnadj && nn_stage_order(f.W)) ? f.W / 2 : 0;
        f.fam = kGroupTiles[f.var].fam == 0 ? 0 : f.nnadj ? 1 : 2;
        const int nw = kFamWaves[f.fam];
        f.ntiles = ((long)e.B + (long)nw * G * f.V - 1) / ((long)nw * G * f.V);
        // Layers per LDS stage: at least one unrolled group of the adjoint loop.
        const int depth = f.W <= 4 ? 4 : 2;
        const size_t per_layer = (size_t)nw * (f.NP / 2) * sizeof(float4);  // K-form: one entry per slot
        int C = (int)(kGroupStage / per_layer);
        if (C > 16) C = 16;
        if (e.tile.chunk >= 1 && (size_t)e.tile.chunk * per_layer <= kGroupStage) C = e.tile.chunk;
        C -= C % depth;
        if (C < depth) C = depth;
        if ((size_t)C * per_layer > kGroupStage) return MESHPROP_ERR_UNSUPPORTED;
        f.C = C;
        if (f.nnadj) f.C = nn_stage_depth(f.W);
        const bool grad = f.what & G_GRAD;
        f.S = (e.B == 0) ? 0 : grad ? (int)std::min<long>(f.ntiles, 2L * g.ncu) : (int)f.ntiles;
        f.E = (long)e.L * (f.NP / 2) * f.np;
        // Tables: caller-owned (built on request) or in the workspace (built every launch).
        size_t tk, tf, ti, td, tt;
        table_layout(e.N, e.L, f.W, tk, tf, ti, td, tt);
        f.tab_kind = tk;
        f.tab_fwd = tf;
        f.tab_inv = ti;
        f.tab_dtdp = td;
        f.own_tables = (e.tables == nullptr);
        f.build = 0;
        if (e.L > 0) {
            if (f.own_tables)
                f.build = B_KIND | ((f.what & G_FWD) ? B_FWD : 0) | (grad ? (B_INV | B_DTDP) : 0);
            else if (e.build_tables)
                f.build = B_KIND | B_FWD | B_INV | B_DTDP;
        }
        f.off_tab = off;
        if (f.own_tables) off = a256(off + tt);
        f.off_partial = off;
        if (grad && e.L > 0 && f.S > 0) off = a256(off + (size_t)f.S * f.E * sizeof(float));
        f.pro_units = f.build ? (long)e.L * ((f.NP / 2 + 255) / 256) : 0;
        const bool assign = (e.flags & MESHPROP_GROUP_DP_ASSIGN) != 0;
        const long red_slots = f.kso ? f.NP / 2 : e.N / 2;  // positions of a stage-order mesh
        f.red_units = (grad && e.L > 0 && (f.S > 0 || assign)) ? ((long)e.L * red_slots * f.np + 63) / 64 : 0;
        g.pro_total += f.pro_units;
        g.red_total += f.red_units;
        // Work items, costed by layers x tile work x tiles per item.
        if (f.S > 0) {
            const long per_item = (f.ntiles + f.S - 1) / f.S;
            const double w = ((f.what & G_FWD) ? 1.0 : 0.0) + (grad ? 3.0 : 0.0);
            const double cost = (double)per_item * ((double)e.L * f.W * f.V * w + 64.0);
            for (int s = 0; s < f.S; ++s) cand.push_back({cost, i, s, f.fam});
        }
    }
    std::stable_sort(cand.begin(), cand.end(), [](const Cand& a, const Cand& b) { return a.cost > b.cost; });
    for (int fam = 0; fam < kGroupFams; ++fam) g.items[fam].clear();
    for (const Cand& c : cand) g.items[c.fam].push_back(make_int2(c.e, c.s));
    // Device plan at the start of the workspace, data after it.
    size_t it[kGroupFams];
    g.header_bytes = item_offsets(g, it);
    g.total = g.header_bytes + off;
    for (int fam = 0; fam < kGroupFams; ++fam) g.grid[fam] = dc ? dc->grid[fam] : g.ncu;
    g.sig = (g.items[0].empty() ? 0 : MESHPROP_GROUP_SIG_NARROW) | (g.items[1].empty() ? 0 : MESHPROP_GROUP_SIG_WIDE) |
            (g.items[2].empty() ? 0 : MESHPROP_GROUP_SIG_WIDE_GENERAL) |
            (g.pro_total > 0 ? MESHPROP_GROUP_SIG_PROLOGUE : 0) | (g.red_total > 0 ? MESHPROP_GROUP_SIG_REDUCE : 0);
    return MESHPROP_OK;
}

void build_image(const meshprop_group& g, unsigned char* ws) {
    const int n = (int)g.ent.size();
    const size_t hdr = a256(sizeof(GHeader));
    const size_t pro = hdr, red = a256(pro + (size_t)(n + 1) * sizeof(long));
    const size_t mesh = a256(red + (size_t)(n + 1) * sizeof(long));
    size_t it[kGroupFams];
    (void)item_offsets(g, it);
    g.image.assign(g.header_bytes, 0);
    unsigned char* im = g.image.data();
    GHeader* H = reinterpret_cast<GHeader*>(im);
    H->nmesh = n;
    H->pro_total = g.pro_total;
    H->red_total = g.red_total;
    H->off_pro = (long)pro;
    H->off_red = (long)red;
    H->off_mesh = (long)mesh;
    for (int fam = 0; fam < kGroupFams; ++fam) {
        H->nitems[fam] = (int)g.items[fam].size();
        H->off_items[fam] = (long)it[fam];
    }
    long* pb = reinterpret_cast<long*>(im + pro);
    long* rb = reinterpret_cast<long*>(im + red);
    GMesh* M = reinterpret_cast<GMesh*>(im + mesh);
    unsigned char* data = ws + g.header_bytes;
    long pacc = 0, racc = 0;
    for (int i = 0; i < n; ++i) {
        const meshprop_group_entry& e = g.ent[i];
        const meshprop_group::Info& f = g.info[i];
        pb[i] = pacc;
        rb[i] = racc;
        pacc += f.pro_units;
        racc += f.red_units;
        GMesh& m = M[i];
        m.what = f.what;
        m.var = f.var;
        m.X = e.X;
        m.N = e.N;
        m.L = e.L;
        m.B = e.B;
        m.fft = e.fft;
        m.LPV = f.LPV;
        m.NP = f.NP;
        m.C = f.C;
        m.nnadj = f.nnadj;
        m.real = f.real;
        m.kso = f.kso;
        m.np = f.np;
        m.ns = f.ns;
        m.ldp = e.ldp;
        m.lds = e.lds;
        m.ldu = e.ldu;
        m.S = f.S;
        m.build = f.build;
        m.assign = (e.flags & MESHPROP_GROUP_DP_ASSIGN) ? 1 : 0;
        m.ntiles = f.ntiles;
        m.E = f.E;
        m.pro_chunks = (f.NP / 2 + 255) / 256;
        m.g0 = e.geom0;
        m.g1 = e.geom1;
        m.p = e.p;
        m.s = (f.ns > 0) ? e.s : nullptr;
        m.dp = e.dp;
        unsigned char* tb = f.own_tables ? data + f.off_tab : (unsigned char*)e.tables;
        m.kind = reinterpret_cast<int*>(tb + f.tab_kind);
        m.rows_fwd = reinterpret_cast<float4*>(tb + f.tab_fwd);
        m.rows_inv = reinterpret_cast<float4*>(tb + f.tab_inv);
        m.dtdp = reinterpret_cast<float2*>(tb + f.tab_dtdp);
        m.partial = reinterpret_cast<float*>(data + f.off_partial);
        m.u_in = (const float2*)e.u_in;
        m.u_out = (float2*)e.u_out;
        m.target = (const float2*)e.target;
        m.g_out = (const float2*)e.dJdu_out;
        m.g_in = (float2*)e.dJdu_in;
        m.u_rec = (e.op == MESHPROP_OP_BACKDIFF) ? (float2*)e.u_in : nullptr;
    }
    pb[n] = pacc;
    rb[n] = racc;
    for (int fam = 0; fam < kGroupFams; ++fam)
        if (!g.items[fam].empty()) std::memcpy(im + it[fam], g.items[fam].data(), g.items[fam].size() * sizeof(int2));
}

int hip_status(hipError_t e) {
    if (e == hipSuccess) return MESHPROP_OK;
    return MESHPROP_ERR_HIP;
}

}  // namespace

extern "C" {

int meshprop_group_create(const meshprop_group_entry* entries, int n, meshprop_group** out) {
    if (!out || n < 0 || (n > 0 && !entries)) return MESHPROP_ERR_ARG;
    *out = nullptr;
    for (int i = 0; i < n; ++i) {
        const int rc = check_entry(entries[i]);
        if (rc) return rc;
    }
    meshprop_group* g = new (std::nothrow) meshprop_group();
    if (!g) return MESHPROP_ERR_ARG;
    g->ent.assign(entries, entries + n);
    int rc = plan_group(*g);
    if (rc) {
        delete g;
        return rc;
    }
    *out = g;
    return MESHPROP_OK;
}

void meshprop_group_destroy(meshprop_group* g) { delete g; }

size_t meshprop_group_workspace_size(const meshprop_group* g) { return g ? g->total : 0; }

size_t meshprop_group_tables_size(int X, int N, int L, int fft, const meshprop_tile* tile) {
    (void)fft;
    if (X < 0 || X > 3 || N <= 0 || (N & 1) || N > 1024 || L < 0) return 0;
    meshprop_tile t = tile ? *tile : meshprop_tile{0, 0, 0, 0};
    size_t k, f, i, d, total;
    table_layout(N, L, group_w(N, t), k, f, i, d, total);
    return total;
}

int meshprop_group_upload(const meshprop_group* g, void* ws, size_t bytes, void* stream) {
    if (!g) return MESHPROP_ERR_ARG;
    if (!ws || bytes < g->total || (((uintptr_t)ws) & 255)) return MESHPROP_ERR_WORKSPACE;
    build_image(*g, (unsigned char*)ws);
    return hip_status(hipMemcpyAsync(ws, g->image.data(), g->header_bytes, hipMemcpyHostToDevice,
                                     (hipStream_t)stream));
}

int meshprop_group_signature(const meshprop_group* g) { return g ? g->sig : 0; }

int meshprop_group_launch(const meshprop_group* g, void* ws, size_t bytes, void* stream) {
    if (!g) return MESHPROP_ERR_ARG;
    if (!ws || bytes < g->total || (((uintptr_t)ws) & 255)) return MESHPROP_ERR_WORKSPACE;
    DeviceGroupConfig* dc = device_config(g->dev);
    if (!dc) return MESHPROP_ERR_ARG;
    clear_pending_hip_error();
    hipStream_t st = (hipStream_t)stream;
    unsigned char* w = (unsigned char*)ws;
    hipError_t e;
    if (g->sig & MESHPROP_GROUP_SIG_PROLOGUE) {
        hipLaunchKernelGGL(k_group_prologue, dim3(4 * g->ncu), dim3(256), 0, st, (const unsigned char*)w);
        if ((e = hipGetLastError()) != hipSuccess) return hip_status(e);
    }
    if (!dc->ok) return MESHPROP_ERR_HIP;
    const bool narrow = g->sig & MESHPROP_GROUP_SIG_NARROW;
    const bool wide = g->sig & MESHPROP_GROUP_SIG_WIDE, wide_gen = g->sig & MESHPROP_GROUP_SIG_WIDE_GENERAL;
    // The W = 16 families (one after the other) run beside family 0 on a second stream (fork /
    // join by events, which a graph capture records as two parallel branches).
    const bool fork = narrow && (wide || wide_gen);
    std::unique_lock<std::mutex> hold(dc->lock, std::defer_lock);
    hipStream_t sw = st;
    if (fork) {
        hold.lock();
        if ((e = hipEventRecord(dc->ev_fork, st)) != hipSuccess) return hip_status(e);
        if ((e = hipStreamWaitEvent(dc->side, dc->ev_fork, 0)) != hipSuccess) return hip_status(e);
        sw = dc->side;
    }
    if (wide) {
        hipLaunchKernelGGL((k_group<1>), dim3(g->grid[1]), dim3(64 * kFamWaves[1]), kGroupLds, sw, w);
        if ((e = hipGetLastError()) != hipSuccess) return hip_status(e);
    }
    if (wide_gen) {
        hipLaunchKernelGGL((k_group<2>), dim3(g->grid[2]), dim3(64 * kFamWaves[2]), kGroupLds, sw, w);
        if ((e = hipGetLastError()) != hipSuccess) return hip_status(e);
    }
    if (fork && (e = hipEventRecord(dc->ev_join, dc->side)) != hipSuccess) return hip_status(e);
    if (narrow) {
        hipLaunchKernelGGL((k_group<0>), dim3(g->grid[0]), dim3(64 * kFamWaves[0]), kGroupLds, st, w);
        if ((e = hipGetLastError()) != hipSuccess) return hip_status(e);
    }
    if (fork && (e = hipStreamWaitEvent(st, dc->ev_join, 0)) != hipSuccess) return hip_status(e);
    if (g->sig & MESHPROP_GROUP_SIG_REDUCE) {
        hipLaunchKernelGGL(k_group_reduce, dim3(8 * g->ncu), dim3(256), 0, st, (const unsigned char*)w);
        if ((e = hipGetLastError()) != hipSuccess) return hip_status(e);
    }
    return MESHPROP_OK;
}

int meshprop_group_entry_info(const meshprop_group* g, int i, meshprop_tile* tile, int* items) {
    if (!g || i < 0 || i >= (int)g->ent.size()) return MESHPROP_ERR_ARG;
    const meshprop_group::Info& f = g->info[i];
    if (tile) {
        tile->w = f.W;
        tile->v = f.V;
        tile->waves = kFamWaves[f.fam];
        tile->chunk = f.C;
    }
    if (items) *items = f.S;
    return MESHPROP_OK;
}

}  // extern "C"
