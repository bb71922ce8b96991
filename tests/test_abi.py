"""C ABI checks that need no GPU: the library loads, exports every symbol the header declares,
and its host-side planning (workspace sizes, status strings, argument validation) behaves."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "meshprop.h")
LIB = os.path.join(ROOT, "meshes_amd", "libmeshprop.so")


def _declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"\b(meshprop_[a-z_0-9]+)\s*\(", txt)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        pytest.fail("libmeshprop.so is not built (run __graft_entry__.build())")
    import meshes_amd._lib as L
    return L.load()


def test_exports_every_declared_symbol(lib):
    names = _declared()
    assert len(names) >= 12
    for n in names:
        assert hasattr(lib, n), n
    import meshes_amd._lib as L
    assert set(L.EXPORTS) == set(names)


def test_status_strings(lib):
    for rc in range(6):
        assert lib.meshprop_status_string(rc)
    assert b"scratch" in lib.meshprop_status_string(5)
    assert b"gfx950" in lib.meshprop_version()


def test_workspace_sizes(lib):
    import meshes_amd._lib as L
    # fwdprop: kind array + one row table (L x NP x 16 B) for complex crossings
    n = L.workspace_size(L.OP_FWDPROP, "mzi", 256, 256, 4096)
    assert n >= 256 * 256 * 16
    # fwddiff needs a second table
    assert L.workspace_size(L.OP_FWDDIFF, "mzi", 256, 256, 4096) >= 2 * 256 * 256 * 16
    # backdiff adds the dT/dp table and the dp partials
    assert L.workspace_size(L.OP_BACKDIFF, "mzi", 256, 256, 4096) > n
    # orth uses real rows (8 B)
    assert L.workspace_size(L.OP_FWDPROP, "orth", 256, 256, 4096) < n
    # invalid sizes give 0
    assert L.workspace_size(L.OP_FWDPROP, "mzi", 255, 4, 4) == 0
    assert L.workspace_size(L.OP_FWDPROP, "mzi", 2048, 4, 4) == 0


def test_fwdgrad_plan_and_arguments_without_gpu(lib):
    """The fused fwd+grad op: a third table beside the adjoint's, unitary (or asserted
    scaled-unitary) crossings only, target required, no checkpoints; all rejected on the host
    before any launch."""
    import meshes_amd._lib as L
    n_b = L.workspace_size(L.OP_BACKDIFF, "mzi", 256, 256, 4096)
    assert L.workspace_size(L.OP_FWDGRAD, "mzi", 256, 256, 4096) >= n_b + 256 * 256 * 16
    assert L.workspace_size(L.OP_FWDGRAD, "orth", 64, 8, 100) > 0
    # generic crossings: sized for the scaled-unitary plan, which the call must assert
    assert L.workspace_size(L.OP_FWDGRAD, "mzi_generic", 64, 8, 100) > 0
    assert L.resolve_tile(L.OP_FWDGRAD, "mzi", 256, 4096) == L.resolve_tile(L.OP_BACKDIFF, "mzi", 256, 4096)
    fake = ctypes.c_void_p(4096)
    geo = (ctypes.c_int * 8)(*([4] * 8))

    def run(X=L.MZI, target=fake, dp=fake, ckpt=None, every=0, op=L.OP_FWDGRAD, flags=0):
        return lib.meshprop_run(op, X, 8, 8, 4, 0, geo, geo, fake, dp, 8, None, 0, fake, fake, fake,
                                None, 8, 0, target, ckpt, every, None, flags, fake, 1 << 30, None)
    assert run(X=L.MZI_GENERIC) == 3
    assert run(X=L.MZI_GENERIC, flags=L.SCALED_UNITARY, target=None) == 1  # accepted up to its arguments
    assert run(target=None) == 1
    assert run(dp=None) == 1
    assert run(ckpt=fake, every=2) == 3
    # the forward-stage flags belong to fwd+grad only
    assert run(op=L.OP_BACKDIFF, flags=L.ONLY_FORWARD) == 1
    assert run(op=L.OP_FWDPROP, target=None, dp=None, flags=L.SKIP_FORWARD) == 1


def test_argument_validation_without_gpu(lib):
    """Bad arguments are rejected on the host before anything is launched."""
    import meshes_amd._lib as L
    fn = lib.meshprop_fwdprop
    fn.argtypes = [ctypes.c_int] * 4 + [ctypes.c_void_p] * 3 + [ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                                                  ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                                                  ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t,
                                                                  ctypes.c_void_p]
    fake = ctypes.c_void_p(256)
    # odd N
    assert fn(L.MZI, 7, 2, 1, fake, fake, fake, 8, None, 0, fake, fake, 7, 0, fake, 1 << 20, None) == 1
    # unsupported mode (only mode 0 exists, gpu/crossing.cu:122-125)
    assert fn(L.MZI, 8, 2, 1, fake, fake, fake, 8, None, 0, fake, fake, 8, 1, fake, 1 << 20, None) == 3
    # missing workspace
    assert fn(L.MZI, 8, 2, 1, fake, fake, fake, 8, None, 0, fake, fake, 8, 0, None, 0, None) == 2
    # ldu < N
    assert fn(L.MZI, 8, 2, 1, fake, fake, fake, 8, None, 0, fake, fake, 6, 0, fake, 1 << 20, None) == 1
    # N beyond 1024
    assert fn(L.MZI, 2048, 2, 1, fake, fake, fake, 2048, None, 0, fake, fake, 2048, 0, fake, 1 << 20, None) == 3
    # B = 0 is a no-op
    assert fn(L.MZI, 8, 2, 0, fake, fake, fake, 8, None, 0, fake, fake, 8, 0, fake, 1 << 20, None) == 0


def test_noalloc_rejects_beyond_scratch_without_gpu(lib):
    """meshprop_run_noalloc checks on the host, before touching a device, that the call fits
    its static scratch: row tables beyond MESHPROP_NOALLOC_SCRATCH_BYTES are refused with
    MESHPROP_ERR_SCRATCH (an adjoint whose partials alone do not fit runs on fewer workgroups
    instead); argument errors come first, as from meshprop_run."""
    import meshes_amd._lib as L
    assert b"96" in open(HEADER, "rb").read().split(b"MESHPROP_NOALLOC_SCRATCH_BYTES")[1][:40]
    fake = ctypes.c_void_p(4096)
    geo = (ctypes.c_int * 8192)(*([4] * 8192))

    def run(op, N, Lyr, B, dp=None, target=None):
        return lib.meshprop_run_noalloc(op, L.MZI, N, Lyr, B, 0, geo, geo, fake, dp, N, None, 0, fake, None, fake,
                                        None, N, 0, target, None, 0, None, 0, None)
    # 8192 layers of N=1024 rows: 128 MiB of forward rows alone
    assert L.workspace_size(L.OP_FWDPROP, "mzi", 1024, 8192, 64) > L.NOALLOC_SCRATCH_BYTES
    assert run(L.OP_FWDPROP, 1024, 8192, 64) == L.ERR_SCRATCH
    assert run(L.OP_FWDGRAD, 1024, 8192, 64, dp=fake, target=fake) == L.ERR_SCRATCH
    assert run(L.OP_FWDPROP, 1023, 8, 64) == 1  # odd N
    assert run(L.OP_FWDPROP, 2048, 8, 64) == 3  # N beyond 1024


def _entry(Lb, **kw):
    e = Lb.GroupEntry()
    fake = 4096
    e.op, e.X, e.N, e.L, e.B, e.fft = Lb.OP_FWDGRAD, Lb.MZI, 64, 64, 100, 0
    e.geom0 = e.geom1 = e.p = e.dp = e.u_in = e.u_out = e.target = fake
    for k, v in kw.items():
        setattr(e, k, v)
    e.ldp = kw.get("ldp", e.N // 2 * 2)
    e.ldu = kw.get("ldu", e.N)
    return e


def test_group_plan_without_gpu(lib):
    """Grouped-launch planning and its argument checks run on the host alone."""
    import meshes_amd._lib as L
    plan = ctypes.c_void_p()
    ents = (L.GroupEntry * 3)(_entry(L), _entry(L, N=512, L=512, B=2048), _entry(L, op=L.OP_FWDPROP, N=1024, L=10))
    assert lib.meshprop_group_create(ents, 3, ctypes.byref(plan)) == 0
    try:
        # tables + partials of each entry fit in the workspace
        assert lib.meshprop_group_workspace_size(plan) > 512 * 512 * 16 * 2
        t, n = L.Tile(), ctypes.c_int()
        assert lib.meshprop_group_entry_info(plan, 1, ctypes.byref(t), ctypes.byref(n)) == 0
        assert (t.w, t.v) == (8, 1) and n.value >= 1
        assert lib.meshprop_group_entry_info(plan, 2, ctypes.byref(t), ctypes.byref(n)) == 0
        assert t.w == 16 and lib.meshprop_group_signature(plan) == 15  # narrow, wide, prologue, reduce
        assert lib.meshprop_group_entry_info(plan, 3, ctypes.byref(t), ctypes.byref(n)) == 1
        # missing or misaligned workspace
        assert lib.meshprop_group_upload(plan, None, 0, None) == 2
        assert lib.meshprop_group_launch(plan, ctypes.c_void_p(4096 + 8), 1 << 40, None) == 2
    finally:
        lib.meshprop_group_destroy(plan)
    bad = [dict(N=63), dict(op=L.OP_FWDDIFF), dict(X=L.MZI_GENERIC), dict(N=2048), dict(ldu=32),
           dict(target=None), dict(dp=None), dict(u_in=None), dict(ldp=10)]
    want = [1, 1, 3, 3, 1, 1, 1, 1, 1]
    for kw, rc in zip(bad, want):
        one = (L.GroupEntry * 1)(_entry(L, **kw))
        assert lib.meshprop_group_create(one, 1, ctypes.byref(plan)) == rc, kw
    # real-field (orth) gradient entries are grouped
    one = (L.GroupEntry * 1)(_entry(L, X=L.ORTH))
    assert lib.meshprop_group_create(one, 1, ctypes.byref(plan)) == 0
    lib.meshprop_group_destroy(plan)
    # forward entries take the lossy crossing; tables size covers kind + two row tables + dT/dp
    one = (L.GroupEntry * 1)(_entry(L, op=L.OP_FWDPROP, X=L.MZI_GENERIC, ldp=64))
    assert lib.meshprop_group_create(one, 1, ctypes.byref(plan)) == 0
    lib.meshprop_group_destroy(plan)
    assert lib.meshprop_group_tables_size(L.MZI, 256, 256, 0, None) >= 256 * 256 * 16 * 2 + 256 * 128 * 64
    assert lib.meshprop_group_tables_size(L.ORTH, 256, 256, 0, None) >= 256 * 256 * 8 * 2 + 256 * 128 * 4
    assert lib.meshprop_group_tables_size(L.ORTH, 2048, 16, 0, None) == 0


def test_tile_table_consulted_without_gpu(lib):
    """A call without a tile takes the measured table's entry for its op, crossing type,
    square/fat batch (fat: B > 2N, gpu/mesh.py:161) and next tabulated N at or above its own
    (gpu/mesh.py:84); an explicit tile overrides it.  Host-side planning only."""
    import re
    import meshes_amd._lib as L
    inc = os.path.join(os.path.dirname(__file__), "..", "meshes_amd", "csrc", "tile_table.inc")
    rows = [tuple(int(x) for x in m.groups()) for m in
            (re.match(r"\{(-?\d+), (-?\d+), (\d+), (\d+), (\d+), (\d+), (\d+)\},", ln.strip()) for ln in open(inc))
            if m]
    assert len(rows) > 100
    names = {0: "mzi", 1: "sym", 2: "orth", 3: "mzi_generic"}
    checked = 0
    for op, X, fat, n0, W, V, waves in rows:
        if op == L.OP_FWDGRAD and X == 3:
            continue  # needs the scaled-unitary flag, which resolve_tile does not take
        for N in {n0, max(2, n0 // 2 + 2)}:
            if N > n0 or (n0 > 8 and N <= n0 // 2):
                continue
            B = 2 * N + 1 if fat else 2 * N
            got = L.resolve_tile(op, names[X], N, B)
            cap = 4 if op == L.OP_FWDPROP or op == L.OP_FWDDIFF else 8  # the plan clamps waves
            assert got[:2] == (W, V), (op, X, fat, n0, N, got)
            if W == 16 and X in (0, 1) and op in (L.OP_BACKDIFF, L.OP_FWDGRAD):
                # the NN-specialised W = 16 adjoint picks its workgroup size by batch
                assert got[2] == (8 if B >= 4 * 256 else 4), (op, X, fat, n0, N, got)
            else:
                assert got[2] == min(waves, cap), (op, X, fat, n0, N, got)
            checked += 1
    assert checked > 100
    # an explicit tile wins over the table
    assert L.resolve_tile(L.OP_FWDPROP, "orth", 512, 4096, (8, 2, 4, 0))[:3] == (8, 2, 4)


def test_wrappers_refuse_other_element_types():
    """The ctypes wrappers hand raw pointers to kernels that read complex64 fields (float32 for
    orthogonal crossings) and float32 parameters: other element types are refused."""
    import torch

    import meshes_amd._lib as L
    f32, c64 = torch.zeros(4, dtype=torch.float32), torch.zeros(4, dtype=torch.complex64)
    L._check_dtypes("mzi", (f32, None), (c64, None, c64))
    L._check_dtypes("orth", (f32,), (f32,))
    with pytest.raises(TypeError):
        L._check_dtypes("mzi", (f32,), (c64.to(torch.complex128),))
    with pytest.raises(TypeError):
        L._check_dtypes("orth", (f32,), (c64,))
    with pytest.raises(TypeError):
        L._check_dtypes("sym", (f32.double(),), (c64,))


def test_w16_adjoint_plans_without_gpu(lib):
    """Unitary NN meshes of N > 512 run the NN-specialised adjoint with one-layer stages; the
    general adjoint flag, FFT geometries and real fields keep the general kernel; the workspace
    covers the general plan's larger partial count."""
    import meshes_amd._lib as L
    for tile in (None, (16, 1, 4, 0), (16, 1, 8, 0)):
        assert L.adjoint_kernel(L.OP_FWDGRAD, "mzi", 1024, 64, 2048, tile=tile) == "nn"
        assert L.resolve_tile(L.OP_BACKDIFF, "mzi", 1024, 2048, tile)[3] == 1
    assert L.adjoint_kernel(L.OP_BACKDIFF, "mzi", 1024, 64, 2048, flags=L.GENERIC_ADJOINT) == "general"
    assert L.adjoint_kernel(L.OP_BACKDIFF, "mzi", 1024, 64, 2048, fft=True) == "general"
    assert L.adjoint_kernel(L.OP_BACKDIFF, "orth", 1024, 64, 2048) == "general"
    a = L.workspace_size(L.OP_BACKDIFF, "mzi", 1024, 64, 2048, False, (16, 1, 8, 0))
    b = L.workspace_size(L.OP_BACKDIFF, "mzi", 1024, 64, 2048, False, (16, 1, 4, 0))
    assert a == b  # eight-wave plan sized for the four-wave general plan the flag may select


def test_forward_kernel_choice_without_gpu(lib):
    """The LDS-staged forward serves wide NN meshes with at least two waves per SIMD (this
    host-side plan assumes the 256 CUs of an MI355X when no device is present)."""
    import meshes_amd._lib as L
    assert L.forward_kernel(L.OP_FWDPROP, "mzi", 1024, 64, 2048) == "nn"
    assert L.forward_kernel(L.OP_FWDGRAD, "mzi", 1024, 64, 2048) == "nn"
    assert L.forward_kernel(L.OP_FWDPROP, "mzi", 1024, 64, 1024) == "general"  # one wave per SIMD
    assert L.forward_kernel(L.OP_FWDPROP, "mzi", 512, 64, 4096) == "general"
    assert L.forward_kernel(L.OP_FWDPROP, "mzi", 1024, 64, 2048, fft=True) == "general"
    assert L.forward_kernel(L.OP_FWDPROP, "mzi", 1024, 64, 2048, flags=L.GENERIC_KERNELS) == "general"
    assert L.forward_kernel(L.OP_BACKDIFF, "mzi", 1024, 64, 2048) == "general"
