// screen.hip -- the phase screen of is_phase meshes: u_w <- exp(i phi_w) u_w on every vector,
// its forward-mode derivative and its adjoint with dJ/dphi (c/dot.cpp:97-143 is the CPU
// semantics; the reference GPU path rejects is_phase, gpu/mesh.py:65); and the waveguide
// permutation (gather) between the layers of a permuted mesh (c/dot.cpp:150-152, 289-291).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "meshprop.h"

namespace {

constexpr int kScreenVectors = 64;  // vectors per block of the adjoint (one dphi partial each)

// Fields are (B, N) rows of complex64 with leading dimension ldu.  Block (x, y): waveguides
// 256*x .. 256*x+255 of the vectors kScreenVectors*y ...
__global__ void __launch_bounds__(256) k_screen(int op, int N, int B, const float* __restrict__ phi,
                                                const float* __restrict__ dphi, float2* u, float2* du, int ldu,
                                                float* __restrict__ partial) {
    const int w = blockIdx.x * 256 + threadIdx.x;
    if (w >= N) return;
    double sd, cd;
    sincos((double)phi[w], &sd, &cd);  // rounded once, as the row tables
    const float s = (float)sd, c = (float)cd;
    const long b0 = (long)blockIdx.y * kScreenVectors;
    const long b1 = (b0 + kScreenVectors < B) ? b0 + kScreenVectors : B;
    if (op == MESHPROP_OP_BACKDIFF) {
        // dJ/dphi_w = Im(sum_b g conj(y)) of the field y and gradient g after the screen; then
        // both go back through the screen (multiplied by exp(-i phi)).
        float acc = 0.f;
        for (long b = b0; b < b1; ++b) {
            const size_t o = (size_t)b * ldu + w;
            const float2 y = u[o], g = du[o];
            acc += g.y * y.x - g.x * y.y;
            u[o] = make_float2(y.x * c + y.y * s, y.y * c - y.x * s);
            du[o] = make_float2(g.x * c + g.y * s, g.y * c - g.x * s);
        }
        partial[(size_t)blockIdx.y * N + w] = acc;
        return;
    }
    const float dph = (op == MESHPROP_OP_FWDDIFF && dphi) ? dphi[w] : 0.f;
    for (long b = b0; b < b1; ++b) {
        const size_t o = (size_t)b * ldu + w;
        const float2 x = u[o];
        const float2 y = make_float2(x.x * c - x.y * s, x.x * s + x.y * c);
        u[o] = y;
        if (op == MESHPROP_OP_FWDDIFF) {
            const float2 d = du ? du[o] : make_float2(0.f, 0.f);
            // d' = exp(i phi) d + i dphi y
            du[o] = make_float2(d.x * c - d.y * s - dph * y.y, d.x * s + d.y * c + dph * y.x);
        }
    }
}

// dphi_w += sum over vector blocks (in block order) of the partials.
__global__ void __launch_bounds__(256) k_screen_reduce(int N, int nblk, const float* __restrict__ partial,
                                                       float* __restrict__ dphi) {
    const int w = blockIdx.x * 256 + threadIdx.x;
    if (w >= N) return;
    float acc = 0.f;
    for (int k = 0; k < nblk; ++k) acc += partial[(size_t)k * N + w];
    dphi[w] += acc;
}

inline int nblocks(int B) { return (B + kScreenVectors - 1) / kScreenVectors; }

// out[b][w] = in[b][perm[w]] for w < N: one thread per output element (coalesced writes; the
// reads of a vector stay within its N contiguous elements).  E: element type (complex64 as
// float2, or float).
template <class E>
__global__ void __launch_bounds__(256) k_permute(int N, int B, const int* __restrict__ perm, const E* __restrict__ in,
                                                 E* __restrict__ out, int ldu) {
    const int w = blockIdx.x * 256 + threadIdx.x;
    if (w >= N) return;
    const int src = perm[w];
    const long b1 = ((long)blockIdx.y + 1) * kScreenVectors < B ? ((long)blockIdx.y + 1) * kScreenVectors : B;
    for (long b = (long)blockIdx.y * kScreenVectors; b < b1; ++b) out[(size_t)b * ldu + w] = in[(size_t)b * ldu + src];
}

}  // namespace

extern "C" {

size_t meshprop_phase_screen_workspace_size(int N, int B) {
    if (N <= 0 || B < 0) return 0;
    return (size_t)nblocks(B) * N * sizeof(float);
}

int meshprop_phase_screen(int op, int N, int B, const float* phi, float* dphi, void* u, void* du, int ldu,
                          void* workspace, size_t workspace_bytes, void* stream) {
    if (op != MESHPROP_OP_FWDPROP && op != MESHPROP_OP_FWDDIFF && op != MESHPROP_OP_BACKDIFF) return MESHPROP_ERR_ARG;
    if (N <= 0 || B < 0 || ldu < N) return MESHPROP_ERR_ARG;
    if (B == 0) return MESHPROP_OK;
    if (!phi || !u) return MESHPROP_ERR_ARG;
    if (op == MESHPROP_OP_FWDDIFF && !du) return MESHPROP_ERR_ARG;
    if (op == MESHPROP_OP_BACKDIFF) {
        if (!du || !dphi) return MESHPROP_ERR_ARG;
        if (!workspace || workspace_bytes < meshprop_phase_screen_workspace_size(N, B)) return MESHPROP_ERR_WORKSPACE;
    }
    (void)hipGetLastError();  // clear an error left by earlier work: the checks below report this call's
    hipStream_t st = (hipStream_t)stream;
    const dim3 grid((N + 255) / 256, nblocks(B));
    hipLaunchKernelGGL(k_screen, grid, dim3(256), 0, st, op, N, B, phi, dphi, (float2*)u, (float2*)du, ldu,
                       (float*)workspace);
    if (hipGetLastError() != hipSuccess) return MESHPROP_ERR_HIP;
    if (op == MESHPROP_OP_BACKDIFF) {
        hipLaunchKernelGGL(k_screen_reduce, dim3((N + 255) / 256), dim3(256), 0, st, N, nblocks(B),
                           (const float*)workspace, dphi);
        if (hipGetLastError() != hipSuccess) return MESHPROP_ERR_HIP;
    }
    return MESHPROP_OK;
}

int meshprop_permute(int N, int B, const int* perm, const void* u_in, void* u_out, int ldu, int elem_bytes,
                     void* stream) {
    if (N <= 0 || B < 0 || ldu < N || (elem_bytes != 8 && elem_bytes != 4)) return MESHPROP_ERR_ARG;
    if (B == 0) return MESHPROP_OK;
    if (!perm || !u_in || !u_out || u_in == u_out) return MESHPROP_ERR_ARG;
    (void)hipGetLastError();
    hipStream_t st = (hipStream_t)stream;
    const dim3 grid((N + 255) / 256, nblocks(B));
    if (elem_bytes == 8)
        hipLaunchKernelGGL(k_permute<float2>, grid, dim3(256), 0, st, N, B, perm, (const float2*)u_in, (float2*)u_out,
                           ldu);
    else
        hipLaunchKernelGGL(k_permute<float>, grid, dim3(256), 0, st, N, B, perm, (const float*)u_in, (float*)u_out,
                           ldu);
    if (hipGetLastError() != hipSuccess) return MESHPROP_ERR_HIP;
    return MESHPROP_OK;
}

}  // extern "C"
