/*
 * meshprop.h -- C ABI of the MI355X (gfx950) mesh-propagation library libmeshprop.so.
 *
 * Batched propagation of complex (or, for orthogonal crossings, real) field vectors
 * through L layers of 2x2 photonic crossings, in three directions: forward (fwdprop),
 * forward-mode differentiation (fwddiff) and reverse-mode differentiation (backdiff).
 *
 * Conventions shared with the CUDA kernels these entry points replace:
 *   - N modes (even), L layers, B vectors.  Fields are column-major (N, B) arrays with
 *     leading dimension ldu >= N: vector b starts at u + b*ldu (gpu/mesh.py:150-155).
 *     Element type is complex64 (interleaved float re/im) for MESHPROP_MZI, MESHPROP_SYM
 *     and MESHPROP_MZI_GENERIC, and float32 for MESHPROP_ORTH (gpu/mesh.py:73).
 *   - Parameters use the slot layout p[l*ldp + m*n_p + k], s[l*lds + m*n_s + k] with
 *     ldp = N/2*n_p, lds = N/2*n_s (gpu/mesh.py:154-155, gpu/gmem.cu:236-240).
 *     n_p = 2,2,1,2 and n_s = 2,1,0,14 for MZI, SYM, ORTH, MZI_GENERIC.
 *   - NN geometry: lens[l], shifts[l]; slot m of layer l couples waveguides
 *     (2m + shifts[l]%2, 2m + shifts[l]%2 + 1) and is active iff
 *     shifts[l]/2 <= m < shifts[l]/2 + lens[l]; inactive slots are the identity.
 *     FFT geometry (the *_ft entry points): strides[l]; slot c = blk*S + j, S = 2^strides[l],
 *     couples (2*S*blk + j, 2*S*blk + j + S) (gpu/fwdprop.cu:48-93, gpu/test.py:73-77).
 *   - Gradient convention: dJdu = dJ/dRe(u) + i dJ/dIm(u); dp += Re sum_ij dJdT_ij dT_ij/dp
 *     with dJdT_ij = sum_b conj(dJdu_i) u_j (gpu/crossing.cu:150-203, 285-302).
 *
 * Every pointer argument except `stream` is a device pointer.  Geometry arrays
 * (lens, shifts, strides) are int32 device arrays of length L, as the reference kernels
 * take them.  Calls are asynchronous on `stream` (a hipStream_t; NULL = default stream),
 * never synchronise and never allocate: the caller provides a device workspace of at
 * least meshprop_workspace_size(...) bytes.  Nullable pointers: s (ideal splitters),
 * du_in (zero tangent), du_out, dJdu_in.  backdiff supports u_in == u_out and
 * dJdu_in == dJdu_out (in place), as gpu/mesh.py:228-231 calls it.  `dp` of backdiff is
 * accumulated into (gpu/mesh.py:214 zero-fills it first).  Each entry point returns a
 * meshprop_status code.
 */
#ifndef MESHPROP_H
#define MESHPROP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum meshprop_crossing {
    MESHPROP_MZI = 0,          /* T(theta, phi; alpha, beta): crossing.py:216-221, gpu/crossing.cu:36-84 */
    MESHPROP_SYM = 1,          /* T(theta, phi; alpha): crossing.py:377-381, gpu/crossing.cu:86-126 */
    MESHPROP_ORTH = 2,         /* real T(theta): gpu/crossing.cu:128-148 */
    MESHPROP_MZI_GENERIC = 3   /* lossy MZI with 14 imperfection values: crossing.py:444-491 */
};

enum meshprop_status {
    MESHPROP_OK = 0,
    MESHPROP_ERR_ARG = 1,          /* bad size, stride or null pointer */
    MESHPROP_ERR_WORKSPACE = 2,    /* workspace missing or smaller than meshprop_workspace_size */
    MESHPROP_ERR_UNSUPPORTED = 3,  /* mode != 0, N beyond 1024, FFT stride that does not tile N */
    MESHPROP_ERR_HIP = 4,          /* a HIP launch failed; see meshprop_last_hip_error() */
    MESHPROP_ERR_SCRATCH = 5       /* meshprop_run_noalloc: the call needs more than the static scratch */
};

enum meshprop_op {
    MESHPROP_OP_FWDPROP = 0,
    MESHPROP_OP_FWDDIFF = 1,
    MESHPROP_OP_BACKDIFF = 2,
    MESHPROP_OP_FWDGRAD = 3    /* fwdprop, then backdiff of its own output (meshprop_run, group entries) */
};

/* Optional tile choice.  Zero fields mean "library default".  w: waveguides per lane
 * (2, 4, 8, 16); v: vectors per lane (1, 2, 4, 8); waves: wave64s per workgroup (1..4, or 8
 * for the pipelined unitary adjoints); chunk: layers per LDS reduction stage of backdiff. */
typedef struct meshprop_tile {
    int w, v, waves, chunk;
} meshprop_tile;

/* Bytes of device workspace the given call needs (0 is possible), with any flags of
 * meshprop_run. */
size_t meshprop_workspace_size(int op, int X, int N, int L, int B, int fft, const meshprop_tile *tile);

/* Replaces fwdprop_N{64K}_{mzi,sym,orth} (gpu/fwdprop.cu:25-27; launched at gpu/mesh.py:158-163). */
int meshprop_fwdprop(int X, int N, int L, int B, const int *lens, const int *shifts,
                     const float *p, int ldp, const float *s, int lds,
                     const void *u_in, void *u_out, int ldu, int mode,
                     void *workspace, size_t workspace_bytes, void *stream);

/* Replaces fwdprop_ft_N{64K}_{X} (gpu/consts.cuh:31-33 argument list with strides; gpu/test.py:48-56). */
int meshprop_fwdprop_ft(int X, int N, int L, int B, const int *strides,
                        const float *p, int ldp, const float *s, int lds,
                        const void *u_in, void *u_out, int ldu, int mode,
                        void *workspace, size_t workspace_bytes, void *stream);

/* Replaces fwddiff_N{64K}_{X} (gpu/fwddiff.cu:24-27; launched at gpu/mesh.py:165-174).
 * dp is the tangent direction in parameter space (slot layout, same ldp as p). */
int meshprop_fwddiff(int X, int N, int L, int B, const int *lens, const int *shifts,
                     const float *p, const float *dp, int ldp, const float *s, int lds,
                     const void *u_in, const void *du_in, void *u_out, void *du_out, int ldu, int mode,
                     void *workspace, size_t workspace_bytes, void *stream);

/* Replaces fwddiff_ft_N{64K}_{X}. */
int meshprop_fwddiff_ft(int X, int N, int L, int B, const int *strides,
                        const float *p, const float *dp, int ldp, const float *s, int lds,
                        const void *u_in, const void *du_in, void *u_out, void *du_out, int ldu, int mode,
                        void *workspace, size_t workspace_bytes, void *stream);

/* Replaces backdiff_N{64K}_{X} (gpu/backdiff.cu:24-27; launched at gpu/mesh.py:226-231).
 * Reconstructs u_in from u_out by inverse propagation and accumulates dp.  For
 * MESHPROP_MZI_GENERIC (non-unitary) the reconstruction uses T^-1 and the adjoint T^H. */
int meshprop_backdiff(int X, int N, int L, int B, const int *lens, const int *shifts,
                      const float *p, float *dp, int ldp, const float *s, int lds,
                      const void *u_out, const void *dJdu_out, void *u_in, void *dJdu_in, int ldu, int mode,
                      void *workspace, size_t workspace_bytes, void *stream);

/* Replaces backdiff_ft_N{64K}_{X}. */
int meshprop_backdiff_ft(int X, int N, int L, int B, const int *strides,
                         const float *p, float *dp, int ldp, const float *s, int lds,
                         const void *u_out, const void *dJdu_out, void *u_in, void *dJdu_in, int ldu, int mode,
                         void *workspace, size_t workspace_bytes, void *stream);

/*
 * Extended entry point behind all of the above (no reference counterpart: the reference
 * has one kernel per size and geometry).  Adds a tile override, stage flags and two forms
 * the reference kernels do not have:
 *   - target (backdiff only, nullable): if set, dJdu_out is ignored and the loss gradient
 *     of J = sum |u_out - target|^2, dJdu_out = 2 (u_out - target), is formed in registers
 *     at load time (the loss of gpu/test.py:131, fused into the adjoint's load);
 *   - ckpt/ckpt_every (nullable / 0): fwdprop writes the field entering every layer
 *     l % ckpt_every == 0 into ckpt[(l/ckpt_every)*B*ldu ...]; backdiff reloads it there
 *     instead of trusting the reconstruction (checkpointed adjoint for lossy meshes).
 * fft selects the strides geometry (geom0 = strides, geom1 unused) or NN (geom0 = lens,
 * geom1 = shifts).  dp/du_in/du_out are used by fwddiff; dp/dJdu_out/dJdu_in by backdiff.
 * op = MESHPROP_OP_FWDGRAD (unitary crossings; target required, ckpt unsupported) is the
 * whole gradient step of J = sum |u_out - target|^2 (gpu/mesh.py:208-234 with the loss of
 * gpu/test.py:131) in one call: one prologue builds the forward rows, the inverse rows and
 * dT/dp together; fwdprop takes u_a = u_in to u_b = u_out; the adjoint of that output
 * accumulates dp (or writes it, MESHPROP_DP_ASSIGN) and, if du_b is set, writes dJdu_in
 * there.  The tile is the adjoint's; workspace from meshprop_workspace_size(MESHPROP_OP_FWDGRAD, ...).
 */
int meshprop_run(int op, int X, int N, int L, int B, int fft, const int *geom0, const int *geom1,
                 const float *p, float *dp, int ldp, const float *s, int lds,
                 const void *u_a, const void *du_a, void *u_b, void *du_b, int ldu, int mode,
                 const void *target, void *ckpt, int ckpt_every, const meshprop_tile *tile, int flags,
                 void *workspace, size_t workspace_bytes, void *stream);

/*
 * meshprop_run without a workspace argument.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments, semantics and results (within
 * FP32 rounding; bitwise equal where the partials fit, see below) as meshprop_run, with the
 * workspace replaced by a file-scope static device array of MESHPROP_NOALLOC_SCRATCH_BYTES
 * (96 MiB) that every device the library runs on holds until the process ends.  The call
 * allocates nothing (no hipMalloc, hipFree or allocating library routine).  If the row tables
 * alone exceed the scratch it returns MESHPROP_ERR_SCRATCH; if the adjoint's per-workgroup dp
 * partials do not fit, the adjoint runs on fewer workgroups (each walking more vector tiles)
 * so that they do: the summation order then differs from meshprop_run's, not its accuracy.
 * Every call on a device shares that device's scratch and leaves its contents there (the
 * stage flags MESHPROP_SKIP_PROLOGUE / MESHPROP_ONLY_REDUCE rely on that): calls that overlap
 * on one device (different streams or host threads) race and give wrong results.  The first
 * call on a device looks up the scratch address there (hipGetSymbolAddress, no allocation).
 */
#define MESHPROP_NOALLOC_SCRATCH_BYTES ((size_t)96 << 20)
int meshprop_run_noalloc(int op, int X, int N, int L, int B, int fft, const int *geom0, const int *geom1,
                         const float *p, float *dp, int ldp, const float *s, int lds,
                         const void *u_a, const void *du_a, void *u_b, void *du_b, int ldu, int mode,
                         const void *target, void *ckpt, int ckpt_every, const meshprop_tile *tile, int flags,
                         void *stream);

/* flags of meshprop_run.  A call is three stages: (1) the prologue that builds the row
 * tables from p, s (and the tangent dp for fwddiff) into the workspace, (2) the
 * propagation kernel, (3) for backdiff, the reduction of dp partials into dp.
 * MESHPROP_ONLY_PROLOGUE runs (1) only; MESHPROP_SKIP_PROLOGUE reuses tables a previous
 * call with the same op, crossing, sizes, geometry, parameters and tile left in this
 * workspace (parameters held resident between calls); MESHPROP_SKIP_REDUCE leaves the
 * partials in the workspace and MESHPROP_ONLY_REDUCE runs (3) alone.  For
 * MESHPROP_OP_FWDGRAD stage (2) is the forward, then the adjoint: MESHPROP_ONLY_FORWARD and
 * MESHPROP_SKIP_FORWARD split it (other ops reject both flags). */
enum meshprop_flags {
    MESHPROP_ONLY_PROLOGUE = 1,
    MESHPROP_SKIP_PROLOGUE = 2,
    MESHPROP_SKIP_REDUCE = 4,
    MESHPROP_ONLY_REDUCE = 8,
    MESHPROP_DP_ASSIGN = 16,   /* backdiff / fwd+grad: dp = gradient instead of dp += gradient */
    MESHPROP_ONLY_FORWARD = 32, /* fwd+grad: run the forward propagation alone */
    MESHPROP_SKIP_FORWARD = 64, /* fwd+grad: u_b already holds the forward output; run the rest */
    /* MESHPROP_MZI_GENERIC backdiff / fwd+grad: the caller asserts that every crossing is a
     * scaled unitary, T = c U (balanced arm gains: g11 = g21 and g12 = g22 in the splitter
     * row).  The adjoint then runs as for unitary crossings on one table of inverse rows, with
     * T^H = |det T| T^-1, and fwd+grad accepts the generic crossing.  The prologue checks the
     * assertion in FP64 per crossing; where it fails, dp and dJ/du_in come out NaN. */
    MESHPROP_SCALED_UNITARY = 128,
    /* Neighbour (NN) geometries: run the general propagation kernels instead of the ones
     * specialised for such meshes -- the forward of fwdprop and fwd+grad, and the unitary
     * adjoint of backdiff and fwd+grad (bitwise equal results when both run the same
     * workgroup shape; for comparison and timing).  MESHPROP_GENERIC_ADJOINT is the earlier
     * name of the same flag. */
    MESHPROP_GENERIC_KERNELS = 256,
    MESHPROP_GENERIC_ADJOINT = MESHPROP_GENERIC_KERNELS
};

/*
 * Grouped launch: many meshes of different width, depth and geometry in one persistent
 * propagation kernel (no reference counterpart: the reference launches one kernel per mesh
 * and call, gpu/mesh.py:158-231).  Work items are (mesh, partial slot) pairs taken from a
 * device-side queue by resident workgroups, longest first; an item runs each of its vector
 * tiles through the whole depth (forward, then the adjoint for gradient entries).
 * A launch is a prologue (row tables of every entry that needs them), the persistent
 * kernel and the dp reduction, all grid-stride kernels whose launch shape depends only on
 * the device and on which of them the plan needs (meshprop_group_signature): every plan
 * with the same signature uploaded into the same workspace replays one captured hipGraph.
 *
 * Entry fields.  op: MESHPROP_OP_FWDPROP (u_in -> u_out), MESHPROP_OP_BACKDIFF (u_out,
 * dJdu_out or target -> dp, dJdu_in and, if u_in is set, the reconstructed input written
 * to u_in) or MESHPROP_OP_FWDGRAD (u_in -> u_out, then the adjoint of that output).
 * Gradient entries need a unitary crossing (MESHPROP_MZI, MESHPROP_SYM, MESHPROP_ORTH);
 * forward entries take any crossing.  MESHPROP_ORTH entries carry float32 fields, the others
 * complex64; one plan may mix them.  target, when
 * set, replaces dJdu_out by 2 (u_out - target).  dp is accumulated into (or, with
 * MESHPROP_GROUP_DP_ASSIGN in flags, overwritten on its N/2 slots per layer).  tile.w, tile.v
 * and tile.chunk pick the tile (0 = default); tile.waves is ignored.  tables, when set,
 * is a caller-owned buffer of meshprop_group_tables_size() bytes holding the entry's
 * row tables across calls (device-resident phases): they are rebuilt from p only when
 * build_tables is nonzero.  Otherwise the tables live in the workspace and are rebuilt on
 * every launch.
 */
typedef struct meshprop_group_entry {
    int op, X, N, L, B, fft;
    const int *geom0, *geom1;
    const float *p;
    int ldp;
    const float *s;
    int lds;
    float *dp;
    const void *u_in;
    void *u_out;
    const void *target;
    const void *dJdu_out;
    void *dJdu_in;
    int ldu;
    meshprop_tile tile;
    void *tables;
    int build_tables;
    int flags;             /* MESHPROP_GROUP_DP_ASSIGN: write dp instead of accumulating */
} meshprop_group_entry;

enum meshprop_group_flags {
    MESHPROP_GROUP_DP_ASSIGN = 1
};

typedef struct meshprop_group meshprop_group; /* host-side plan of a grouped launch */

/* Plans a grouped launch of n entries (copied; pointers are kept, not dereferenced). */
int meshprop_group_create(const meshprop_group_entry *entries, int n, meshprop_group **out);
void meshprop_group_destroy(meshprop_group *g);
/* Device workspace bytes of the plan (descriptors, queue, tables, dp partials). */
size_t meshprop_group_workspace_size(const meshprop_group *g);
/* Bytes of a caller-owned table buffer for one entry of this crossing, size and tile. */
size_t meshprop_group_tables_size(int X, int N, int L, int fft, const meshprop_tile *tile);
/* Writes the plan's descriptors into the workspace (one host-to-device copy on stream).
 * The plan must stay alive until the copy has completed. */
int meshprop_group_upload(const meshprop_group *g, void *workspace, size_t workspace_bytes, void *stream);
/* Launches prologue, persistent kernel and reduction on stream for the descriptors last
 * uploaded into the workspace.  The launch shape depends on the device and on
 * meshprop_group_signature() only. */
int meshprop_group_launch(const meshprop_group *g, void *workspace, size_t workspace_bytes, void *stream);
/* Launch-shape signature (MESHPROP_GROUP_SIG_* bits: which of the prologue, the three
 * persistent kernel families and the reduction the plan launches).  Plans with equal
 * signatures uploaded into the same workspace replay the same captured graph. */
int meshprop_group_signature(const meshprop_group *g);
enum meshprop_group_signature_bits {
    MESHPROP_GROUP_SIG_NARROW = 1,        /* items of meshes with N <= 512 */
    MESHPROP_GROUP_SIG_WIDE = 2,          /* items of NN meshes with N > 512 */
    MESHPROP_GROUP_SIG_PROLOGUE = 4,      /* row tables to build */
    MESHPROP_GROUP_SIG_REDUCE = 8,        /* dp to reduce */
    MESHPROP_GROUP_SIG_WIDE_GENERAL = 16  /* items of other meshes with N > 512 (FFT geometry) */
};
/* Resolved tile and number of work items of entry i. */
int meshprop_group_entry_info(const meshprop_group *g, int i, meshprop_tile *tile, int *items);

/* Phase screen of is_phase meshes: u_w <- exp(i phi_w) u_w on every vector (no reference GPU
 * counterpart: gpu/mesh.py:65 rejects is_phase; the CPU semantics are c/dot.cpp:97-143 and
 * mesh.py:38-41).  In place on complex64 (N, B) fields with leading dimension ldu.
 *   MESHPROP_OP_FWDPROP : u <- exp(i phi) u.
 *   MESHPROP_OP_FWDDIFF : u <- exp(i phi) u;  du <- exp(i phi) du + i dphi (exp(i phi) u).
 *   MESHPROP_OP_BACKDIFF: u, du = field and dJ/du after the screen: dphi += Im(sum_b du conj(u))
 *                         (accumulated, fixed summation order), then u <- exp(-i phi) u and
 *                         du <- exp(-i phi) du.  Needs meshprop_phase_screen_workspace_size bytes. */
int meshprop_phase_screen(int op, int N, int B, const float *phi, float *dphi, void *u, void *du, int ldu,
                          void *workspace, size_t workspace_bytes, void *stream);
size_t meshprop_phase_screen_workspace_size(int N, int B);

/* Waveguide permutation of B vectors (rows of ldu elements): u_out[b][w] = u_in[b][perm[w]]
 * for w < N, perm a device int32[N] permutation (the gather v <- v[p] of mesh.py:47-48, 71-72
 * and c/dot.cpp:150-152, 289-291, applied between layers of a permuted StructuredMeshNetwork;
 * the reference GPU path rejects permutations, mesh.py:573).  elem_bytes: 8 for complex64, 4
 * for float32 fields.  u_in and u_out must not overlap.  Asynchronous on stream. */
int meshprop_permute(int N, int B, const int *perm, const void *u_in, void *u_out, int ldu, int elem_bytes,
                     void *stream);

/* Exact Hessian-vector product along the parameter direction q, by forward-over-reverse
 * differentiation (no reference counterpart: gpu/mesh.py:236-288 takes central differences of
 * two backdiffs).  Inputs: the forward output u_out and its tangent du_out along q (from
 * meshprop_fwddiff with dp = q), and either target (loss sum |u_out - target|^2: dJ/du_out =
 * 2 (u_out - target), tangent 2 du_out) or dJdu_out with its tangent d_dJdu_out along q.
 * Accumulates grad += dJ/dp (nullable) and hvp += (d2J/dp2) q in the slot layout of p.  Unitary
 * crossings: MESHPROP_MZI and MESHPROP_SYM (complex64 fields) with N <= 1024 on NN meshes and
 * N <= 512 on FFT meshes; MESHPROP_ORTH (float32 fields) with N <= 1024 on NN meshes.  Other
 * crossings and sizes return MESHPROP_ERR_UNSUPPORTED.  Workspace: meshprop_hvp_workspace_size
 * bytes (it depends on fft). */
size_t meshprop_hvp_workspace_size(int X, int N, int L, int B, int fft);
int meshprop_hvp(int X, int N, int L, int B, int fft, const int *geom0, const int *geom1,
                 const float *p, const float *q, int ldp, const float *s, int lds,
                 const void *u_out, const void *du_out, const void *dJdu_out, const void *d_dJdu_out,
                 const void *target, float *grad, float *hvp, int ldu,
                 void *workspace, size_t workspace_bytes, void *stream);
/* meshprop_hvp that also writes the input gradient dJ/du_in and its tangent along q into dJdu_in
 * and d_dJdu_in (both or neither; same layout and ldu as the fields): the walk's final state,
 * which an input phase screen's Hessian entries need.  No reference counterpart, as above. */
int meshprop_hvp_grad_in(int X, int N, int L, int B, int fft, const int *geom0, const int *geom1,
                         const float *p, const float *q, int ldp, const float *s, int lds,
                         const void *u_out, const void *du_out, const void *dJdu_out,
                         const void *d_dJdu_out, const void *target, float *grad, float *hvp,
                         void *dJdu_in, void *d_dJdu_in, int ldu,
                         void *workspace, size_t workspace_bytes, void *stream);

/* The tile the library would use for a call (after defaults and clamping). */
int meshprop_resolve_tile(int op, int X, int N, int B, meshprop_tile *out);

/* Which adjoint kernel a backdiff / fwd+grad call with these arguments runs (no reference
 * counterpart: the reference has one backdiff kernel per size): 1 = the one specialised for
 * unitary crossings on NN geometries, 0 = the general one (also for ops without an adjoint),
 * -status for arguments meshprop_run would reject.  fft as in meshprop_run, flags its flags. */
int meshprop_adjoint_kernel(int op, int X, int N, int L, int B, int fft, const meshprop_tile *tile, int flags);

/* Which forward kernel a fwdprop / fwd+grad (or forward-mode, fwddiff) call with these
 * arguments runs (no reference counterpart): 1 = the one specialised for NN geometries with
 * rows staged through LDS, 0 = the general one (also for backdiff), -status for arguments
 * meshprop_run would reject. */
int meshprop_forward_kernel(int op, int X, int N, int L, int B, int fft, const meshprop_tile *tile, int flags);

/* Last HIP error code seen by a failing call on this thread (hipError_t as int).  A HIP error
 * left pending on the thread by earlier, unrelated work is cleared when a call starts launching,
 * so MESHPROP_ERR_HIP reports a failure of that call's own launches. */
int meshprop_last_hip_error(void);

/* Human-readable status. */
const char *meshprop_status_string(int status);

/* Library version string. */
const char *meshprop_version(void);

#ifdef __cplusplus
}
#endif

#endif /* MESHPROP_H */
