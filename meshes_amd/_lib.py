"""ctypes binding of libmeshprop.so (the C ABI declared in include/meshprop.h).

The library is the only compute path: there is no CPU fallback.  If it is missing, every
entry point raises ``MeshpropUnavailable``.  torch is imported before the library is
loaded so that both share the HIP runtime torch already mapped.
"""
from __future__ import annotations

import ctypes
import os

import torch  # noqa: F401  (must precede loading libmeshprop.so: one HIP runtime per process)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("MESHPROP_LIB", os.path.join(_HERE, "libmeshprop.so"))

MZI, SYM, ORTH, MZI_GENERIC = 0, 1, 2, 3
OP_FWDPROP, OP_FWDDIFF, OP_BACKDIFF, OP_FWDGRAD = 0, 1, 2, 3
GROUP_DP_ASSIGN = 1
X_CODE = {"mzi": MZI, "sym": SYM, "orth": ORTH, "mzi_generic": MZI_GENERIC}
N_PHASE = {"mzi": 2, "sym": 2, "orth": 1, "mzi_generic": 2}
N_SPLIT = {"mzi": 2, "sym": 1, "orth": 0, "mzi_generic": 14}

STATUS = {0: "ok", 1: "invalid argument", 2: "workspace missing or too small", 3: "unsupported configuration",
          4: "HIP launch error", 5: "needs more than the static scratch of meshprop_run_noalloc"}
ERR_SCRATCH = 5

# Every symbol include/meshprop.h declares.
EXPORTS = (
    "meshprop_workspace_size", "meshprop_fwdprop", "meshprop_fwdprop_ft", "meshprop_fwddiff",
    "meshprop_fwddiff_ft", "meshprop_backdiff", "meshprop_backdiff_ft", "meshprop_run",
    "meshprop_resolve_tile", "meshprop_last_hip_error", "meshprop_status_string", "meshprop_version",
    "meshprop_group_create", "meshprop_group_destroy", "meshprop_group_workspace_size",
    "meshprop_group_tables_size", "meshprop_group_upload", "meshprop_group_launch", "meshprop_group_signature",
    "meshprop_group_entry_info", "meshprop_phase_screen", "meshprop_phase_screen_workspace_size",
    "meshprop_permute",
    "meshprop_hvp", "meshprop_hvp_grad_in", "meshprop_hvp_workspace_size", "meshprop_run_noalloc", "meshprop_adjoint_kernel", "meshprop_forward_kernel",
)

# Bytes of the static device scratch behind meshprop_run_noalloc (MESHPROP_NOALLOC_SCRATCH_BYTES).
NOALLOC_SCRATCH_BYTES = 96 << 20


class MeshpropUnavailable(RuntimeError):
    pass


class MeshpropError(RuntimeError):
    pass


class Tile(ctypes.Structure):
    _fields_ = [("w", ctypes.c_int), ("v", ctypes.c_int), ("waves", ctypes.c_int), ("chunk", ctypes.c_int)]


class GroupEntry(ctypes.Structure):
    """meshprop_group_entry (include/meshprop.h)."""
    _fields_ = [("op", ctypes.c_int), ("X", ctypes.c_int), ("N", ctypes.c_int), ("L", ctypes.c_int),
                ("B", ctypes.c_int), ("fft", ctypes.c_int), ("geom0", ctypes.c_void_p), ("geom1", ctypes.c_void_p),
                ("p", ctypes.c_void_p), ("ldp", ctypes.c_int), ("s", ctypes.c_void_p), ("lds", ctypes.c_int),
                ("dp", ctypes.c_void_p), ("u_in", ctypes.c_void_p), ("u_out", ctypes.c_void_p),
                ("target", ctypes.c_void_p), ("dJdu_out", ctypes.c_void_p), ("dJdu_in", ctypes.c_void_p),
                ("ldu", ctypes.c_int), ("tile", Tile), ("tables", ctypes.c_void_p), ("build_tables", ctypes.c_int),
                ("flags", ctypes.c_int)]


_lib = None

c_int, c_size_t, c_void_p = ctypes.c_int, ctypes.c_size_t, ctypes.c_void_p


def load():
    """Load libmeshprop.so once; raise MeshpropUnavailable if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise MeshpropUnavailable(
            f"{LIB_PATH} is missing: build it with `make -C meshes_amd/csrc` (or __graft_entry__.build()).")
    lib = ctypes.CDLL(LIB_PATH)
    lib.meshprop_workspace_size.argtypes = [c_int, c_int, c_int, c_int, c_int, c_int, ctypes.POINTER(Tile)]
    lib.meshprop_workspace_size.restype = c_size_t
    lib.meshprop_run.argtypes = ([c_int] * 6 + [c_void_p] * 3 + [c_void_p, c_int, c_void_p, c_int]
                                 + [c_void_p] * 4 + [c_int, c_int] + [c_void_p, c_void_p, c_int]
                                 + [ctypes.POINTER(Tile), c_int, c_void_p, c_size_t, c_void_p])
    lib.meshprop_run.restype = c_int
    lib.meshprop_run_noalloc.argtypes = lib.meshprop_run.argtypes[:24] + [c_void_p]
    lib.meshprop_run_noalloc.restype = c_int
    lib.meshprop_resolve_tile.argtypes = [c_int, c_int, c_int, c_int, ctypes.POINTER(Tile)]
    lib.meshprop_resolve_tile.restype = c_int
    lib.meshprop_adjoint_kernel.argtypes = [c_int, c_int, c_int, c_int, c_int, c_int, ctypes.POINTER(Tile), c_int]
    lib.meshprop_adjoint_kernel.restype = c_int
    lib.meshprop_forward_kernel.argtypes = lib.meshprop_adjoint_kernel.argtypes
    lib.meshprop_forward_kernel.restype = c_int
    lib.meshprop_last_hip_error.restype = c_int
    lib.meshprop_status_string.restype = ctypes.c_char_p
    lib.meshprop_status_string.argtypes = [c_int]
    lib.meshprop_version.restype = ctypes.c_char_p
    vp = c_void_p
    lib.meshprop_group_create.argtypes = [ctypes.POINTER(GroupEntry), c_int, ctypes.POINTER(vp)]
    lib.meshprop_group_create.restype = c_int
    lib.meshprop_group_destroy.argtypes = [vp]
    lib.meshprop_group_destroy.restype = None
    lib.meshprop_group_workspace_size.argtypes = [vp]
    lib.meshprop_group_workspace_size.restype = c_size_t
    lib.meshprop_group_tables_size.argtypes = [c_int, c_int, c_int, c_int, ctypes.POINTER(Tile)]
    lib.meshprop_group_tables_size.restype = c_size_t
    lib.meshprop_group_upload.argtypes = [vp, vp, c_size_t, vp]
    lib.meshprop_group_upload.restype = c_int
    lib.meshprop_group_launch.argtypes = [vp, vp, c_size_t, vp]
    lib.meshprop_group_launch.restype = c_int
    lib.meshprop_group_signature.argtypes = [vp]
    lib.meshprop_group_signature.restype = c_int
    lib.meshprop_group_entry_info.argtypes = [vp, c_int, ctypes.POINTER(Tile), ctypes.POINTER(c_int)]
    lib.meshprop_group_entry_info.restype = c_int
    lib.meshprop_phase_screen.argtypes = [c_int, c_int, c_int, vp, vp, vp, vp, c_int, vp, c_size_t, vp]
    lib.meshprop_phase_screen.restype = c_int
    lib.meshprop_phase_screen_workspace_size.argtypes = [c_int, c_int]
    lib.meshprop_phase_screen_workspace_size.restype = c_size_t
    lib.meshprop_permute.argtypes = [c_int, c_int, vp, vp, vp, c_int, c_int, vp]
    lib.meshprop_permute.restype = c_int
    lib.meshprop_hvp_workspace_size.argtypes = [c_int] * 5
    lib.meshprop_hvp_workspace_size.restype = c_size_t
    lib.meshprop_hvp.argtypes = ([c_int] * 5 + [vp] * 4 + [c_int, vp, c_int] + [vp] * 7 + [c_int, vp, c_size_t, vp])
    lib.meshprop_hvp.restype = c_int
    lib.meshprop_hvp_grad_in.argtypes = ([c_int] * 5 + [vp] * 4 + [c_int, vp, c_int] + [vp] * 9
                                         + [c_int, vp, c_size_t, vp])
    lib.meshprop_hvp_grad_in.restype = c_int
    _lib = lib
    return lib


def check(rc: int, what: str = "meshprop"):
    if rc != 0:
        lib = load()
        msg = lib.meshprop_status_string(rc).decode()
        extra = f" (hipError {lib.meshprop_last_hip_error()})" if rc == 4 else ""
        raise MeshpropError(f"{what}: {msg}{extra}")


def _ptr(t):
    if t is None:
        return None
    if t.is_conj() or t.is_neg():
        raise ValueError("lazy conjugate / negative tensor views must be resolved before a kernel reads them")
    return ctypes.c_void_p(t.data_ptr())


def _check_dtypes(X, params, fields):
    """The kernels read float32 parameters and complex64 fields (float32 for orthogonal
    crossings) through raw pointers: any other element type is refused here."""
    xc = X_CODE[X] if isinstance(X, str) else X
    want = torch.float32 if xc == X_CODE["orth"] else torch.complex64
    for t in params:
        if t is not None and t.dtype != torch.float32:
            raise TypeError(f"parameters must be float32, got {t.dtype}")
    for t in fields:
        if t is not None and t.dtype != want:
            raise TypeError(f"fields must be {want}, got {t.dtype}")


def tile_struct(tile):
    if tile is None:
        return None
    if isinstance(tile, Tile):
        return ctypes.byref(tile)
    w, v, waves, chunk = (list(tile) + [0, 0, 0, 0])[:4]
    return ctypes.byref(Tile(int(w), int(v), int(waves), int(chunk)))


def workspace_size(op, X, N, L, B, fft=False, tile=None) -> int:
    lib = load()
    tp = None
    if tile is not None:
        w, v, waves, chunk = (list(tile) + [0, 0, 0, 0])[:4]
        tp = ctypes.pointer(Tile(int(w), int(v), int(waves), int(chunk)))
    return int(lib.meshprop_workspace_size(op, X_CODE[X] if isinstance(X, str) else X, N, L, B, int(fft), tp))


def resolve_tile(op, X, N, B, tile=None):
    lib = load()
    w, v, waves, chunk = (list(tile or []) + [0, 0, 0, 0])[:4]
    t = Tile(int(w), int(v), int(waves), int(chunk))
    check(lib.meshprop_resolve_tile(op, X_CODE[X] if isinstance(X, str) else X, N, B, ctypes.byref(t)),
          "resolve_tile")
    return (t.w, t.v, t.waves, t.chunk)


def forward_kernel(op, X, N, L, B, fft=False, tile=None, flags=0) -> str:
    """Which forward kernel a fwdprop / fwd+grad / fwddiff call runs: "nn" (NN geometry, rows
    staged through LDS) or "general"."""
    lib = load()
    rc = lib.meshprop_forward_kernel(op, X_CODE[X] if isinstance(X, str) else X, N, L, B, int(fft),
                                     tile_struct(tile), int(flags))
    if rc < 0:
        check(-rc, "forward_kernel")
    return "nn" if rc == 1 else "general"


def adjoint_kernel(op, X, N, L, B, fft=False, tile=None, flags=0) -> str:
    """Which adjoint kernel a backdiff / fwd+grad call runs: "nn" (specialised for unitary
    crossings on NN geometries) or "general"."""
    lib = load()
    rc = lib.meshprop_adjoint_kernel(op, X_CODE[X] if isinstance(X, str) else X, N, L, B, int(fft),
                                     tile_struct(tile), int(flags))
    if rc < 0:
        check(-rc, "adjoint_kernel")
    return "nn" if rc == 1 else "general"


ONLY_PROLOGUE, SKIP_PROLOGUE, SKIP_REDUCE, ONLY_REDUCE, DP_ASSIGN = 1, 2, 4, 8, 16
ONLY_FORWARD, SKIP_FORWARD = 32, 64
SCALED_UNITARY = 128
GENERIC_KERNELS = GENERIC_ADJOINT = 256  # general kernels instead of the NN-specialised ones


def run(op, X, N, L, B, fft, g0, g1, p, dp, ldp, s, lds, u_a, du_a, u_b, du_b, ldu, target=None, ckpt=None,
        ckpt_every=0, tile=None, workspace=None, ws_bytes=0, stream=None, flags=0):
    """Thin call of meshprop_run on torch tensors (device pointers) and the current stream."""
    lib = load()
    _check_dtypes(X, (p, dp, s), (u_a, du_a, u_b, du_b, target, ckpt))
    if stream is None:
        stream = torch.cuda.current_stream().cuda_stream
    rc = lib.meshprop_run(op, X_CODE[X] if isinstance(X, str) else X, N, L, B, int(fft), _ptr(g0), _ptr(g1),
                          _ptr(p), _ptr(dp), ldp, _ptr(s), lds, _ptr(u_a), _ptr(du_a), _ptr(u_b), _ptr(du_b), ldu,
                          0, _ptr(target), _ptr(ckpt), int(ckpt_every), tile_struct(tile), int(flags), _ptr(workspace),
                          ctypes.c_size_t(ws_bytes), ctypes.c_void_p(stream))
    check(rc, ["fwdprop", "fwddiff", "backdiff", "fwd+grad"][op])


def run_noalloc(op, X, N, L, B, fft, g0, g1, p, dp, ldp, s, lds, u_a, du_a, u_b, du_b, ldu, target=None,
                ckpt=None, ckpt_every=0, tile=None, stream=None, flags=0):
    """meshprop_run_noalloc: meshprop_run with the library's static device scratch in place of a
    workspace (nothing is allocated; calls that overlap on one device race)."""
    lib = load()
    _check_dtypes(X, (p, dp, s), (u_a, du_a, u_b, du_b, target, ckpt))
    if stream is None:
        stream = torch.cuda.current_stream().cuda_stream
    rc = lib.meshprop_run_noalloc(op, X_CODE[X] if isinstance(X, str) else X, N, L, B, int(fft), _ptr(g0), _ptr(g1),
                                  _ptr(p), _ptr(dp), ldp, _ptr(s), lds, _ptr(u_a), _ptr(du_a), _ptr(u_b), _ptr(du_b),
                                  ldu, 0, _ptr(target), _ptr(ckpt), int(ckpt_every), tile_struct(tile), int(flags),
                                  ctypes.c_void_p(stream))
    check(rc, ["fwdprop", "fwddiff", "backdiff", "fwd+grad"][op] + " (no-alloc)")


def phase_screen(op, N, B, phi, dphi, u, du, ldu, workspace=None, ws_bytes=0, stream=None):
    """meshprop_phase_screen on torch tensors (in place on u, du)."""
    lib = load()
    if stream is None:
        stream = torch.cuda.current_stream().cuda_stream
    rc = lib.meshprop_phase_screen(op, N, B, _ptr(phi), _ptr(dphi), _ptr(u), _ptr(du), ldu, _ptr(workspace),
                                   ctypes.c_size_t(ws_bytes), ctypes.c_void_p(stream))
    check(rc, "phase_screen")


def permute(N, B, perm, u_in, u_out, ldu, stream=None):
    """meshprop_permute on torch tensors: u_out[b, w] = u_in[b, perm[w]] (perm int32 on the device)."""
    lib = load()
    if stream is None:
        stream = torch.cuda.current_stream().cuda_stream
    if perm.dtype != torch.int32 or u_in.dtype != u_out.dtype:
        raise TypeError("perm must be int32 and both fields of one element type")
    eb = 8 if u_in.dtype == torch.complex64 else 4 if u_in.dtype == torch.float32 else 0
    if not eb:
        raise TypeError(f"fields must be complex64 or float32, got {u_in.dtype}")
    check(lib.meshprop_permute(N, B, _ptr(perm), _ptr(u_in), _ptr(u_out), ldu, eb, ctypes.c_void_p(stream)),
          "permute")


def phase_screen_workspace_size(N, B) -> int:
    return int(load().meshprop_phase_screen_workspace_size(N, B))


def hvp_workspace_size(X, N, L, B, fft=False) -> int:
    return int(load().meshprop_hvp_workspace_size(X_CODE[X] if isinstance(X, str) else X, N, L, B, int(fft)))


def hvp(X, N, L, B, fft, g0, g1, p, q, ldp, s, lds, u_out, du_out, dJdu_out, d_dJdu_out, target, grad, hv, ldu,
        workspace, ws_bytes, stream=None, dJdu_in=None, d_dJdu_in=None):
    """meshprop_hvp on torch tensors; with dJdu_in and d_dJdu_in, meshprop_hvp_grad_in (which also
    writes the input gradient and its tangent)."""
    lib = load()
    _check_dtypes(X, (p, q, s, grad, hv), (u_out, du_out, dJdu_out, d_dJdu_out, target, dJdu_in, d_dJdu_in))
    if stream is None:
        stream = torch.cuda.current_stream().cuda_stream
    if dJdu_in is not None or d_dJdu_in is not None:
        rc = lib.meshprop_hvp_grad_in(X_CODE[X] if isinstance(X, str) else X, N, L, B, int(fft), _ptr(g0), _ptr(g1),
                                      _ptr(p), _ptr(q), ldp, _ptr(s), lds, _ptr(u_out), _ptr(du_out), _ptr(dJdu_out),
                                      _ptr(d_dJdu_out), _ptr(target), _ptr(grad), _ptr(hv), _ptr(dJdu_in),
                                      _ptr(d_dJdu_in), ldu, _ptr(workspace), ctypes.c_size_t(ws_bytes),
                                      ctypes.c_void_p(stream))
        check(rc, "hvp_grad_in")
        return
    rc = lib.meshprop_hvp(X_CODE[X] if isinstance(X, str) else X, N, L, B, int(fft), _ptr(g0), _ptr(g1), _ptr(p),
                          _ptr(q), ldp, _ptr(s), lds, _ptr(u_out), _ptr(du_out), _ptr(dJdu_out), _ptr(d_dJdu_out),
                          _ptr(target), _ptr(grad), _ptr(hv), ldu, _ptr(workspace), ctypes.c_size_t(ws_bytes),
                          ctypes.c_void_p(stream))
    check(rc, "hvp")
