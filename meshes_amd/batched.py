"""Grouped propagation: many meshes of uneven width, depth and topology in one persistent launch.

A ``MeshGroup`` holds a list of ``GroupJob`` (one mesh and its batch each) and runs all of them
with the library's grouped launch (``meshprop_group_*`` in include/meshprop.h): a prologue that
builds every mesh's row tables, ONE persistent propagation kernel whose resident workgroups take
(mesh, partial slot) items from a device queue longest-first, and the dp reduction.  The plan
lives in the workspace, so the three launches can be captured once as a hipGraph
(``MeshGroup.capture``) and replayed for every later plan uploaded into the same workspace.

Jobs: ``op='fwd'`` computes y = U x; ``op='grad'`` computes y = U x and then the gradient of
J = sum |y - target|^2 (or of a given dJ/dy) with respect to the phases (dp) and, optionally, the
input (gx); ``op='backdiff'`` takes y and dJ/dy (or target) as the reference's backdiff does
(gpu/mesh.py:196-231).  Fields are (B, N) tensors, one vector per row (the (N, B) Fortran arrays
of the reference, transposed).  Gradient jobs need unitary crossings (mzi, sym).

Meshes with a phase screen (``is_phase``: exp(i phi) at the input or output, mesh.py:38-41,
c/dot.cpp:97-143) run their crossings in the group and the screen as launches of the screen
kernel around it (``ScreenPlan``): the screened input or the target seen through the screen
before the group launch, the screened output and the screen's adjoint (dJ/dphi into the last N
entries of dp) after it.  These launches use buffers allocated with the job, so a captured graph
records them too.

The per-mesh reference path, one plain launch sequence per mesh (gpu/mesh.py:158-231, as the
reference runs each mesh), is ``run_plain``.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import torch

from . import _lib
from .mesh import MeshNetworkGPU

OPS = {"fwd": _lib.OP_FWDPROP, "backdiff": _lib.OP_BACKDIFF, "grad": _lib.OP_FWDGRAD}


@dataclass
class GroupJob:
    mesh: MeshNetworkGPU
    op: str
    x: Optional[torch.Tensor] = None        # (B, N) input (fwd, grad); reconstructed input (backdiff, optional)
    y: Optional[torch.Tensor] = None        # (B, N) output (allocated if None)
    target: Optional[torch.Tensor] = None   # (B, N): dJ/dy = 2 (y - target)
    dy: Optional[torch.Tensor] = None       # (B, N) dJ/dy when target is None
    gx: Optional[torch.Tensor] = None       # (B, N) dJ/dx output (optional)
    dp: Optional[torch.Tensor] = None       # (n_phase,) float32 gradient (allocated if None)
    accumulate: bool = False                # dp += gradient instead of dp = gradient
    tile: Optional[Sequence[int]] = None    # (w, v, waves, chunk); None = library default
    tables: Optional[torch.Tensor] = None   # resident row tables (uint8), see tables_for()
    build_tables: bool = True
    B: int = field(init=False, default=0)

    def __post_init__(self):
        assert self.op in OPS, self.op
        m = self.mesh
        if self.op != "fwd":
            assert m.X in ("mzi", "sym", "orth"), "gradient jobs need unitary crossings"
        ref = self.x if self.op != "backdiff" else self.y
        assert ref is not None and ref.ndim == 2 and ref.shape[1] == m.N
        self.B = int(ref.shape[0])
        dev = m.device
        for name in ("x", "y", "target", "dy", "gx"):
            t = getattr(self, name)
            if t is not None:
                assert t.shape == (self.B, m.N) and t.dtype == m.dtype and t.is_contiguous(), name
                assert not (t.is_conj() or t.is_neg()), f"{name}: resolve lazy conjugate/negative views first"
                assert t.device == dev, name
        if self.y is None:
            self.y = torch.empty((self.B, m.N), dtype=m.dtype, device=dev)
        if self.op != "fwd":
            assert self.target is not None or self.dy is not None, "gradient jobs need target or dy"
            if self.dp is None:
                self.dp = torch.zeros(m.n_phase, dtype=torch.float32, device=dev)
        self.screen = ScreenPlan(m, self.op, self.B, x=self.x, y=self.y, target=self.target, dy=self.dy,
                                 gx=self.gx, dp=self.dp, accumulate=self.accumulate) if m.is_phase else None


class ScreenPlan:
    """The phase screen of one grouped job: entry buffers that replace the job's (``sub``) and the
    screen launches before (``pre``) and after (``post``) the group launch.  With S = diag(exp(i
    phi)) and the mesh's crossings M:

    * fwd: in -- the entry reads S x; out -- y <- S y after the launch.
    * grad (J = |y - target|^2 or a given dJ/dy): in -- the entry runs on S x and returns its input
      gradient, whose pass back through the screen's adjoint gives dJ/dphi (and dJ/dx); out -- the
      entry runs against S^H target (or S^H dJ/dy), then y <- S y and the screen's adjoint on
      (y, dJ/dy) gives dJ/dphi.
    * backdiff (y, dJ/dy given): out -- the screen's adjoint maps (y, dJ/dy) to the mesh's output
      and its gradient before the launch; in -- the entry's reconstructed input and input gradient
      go back through the screen's adjoint after it.
    """

    def __init__(self, mesh, op, B, x=None, y=None, target=None, dy=None, gx=None, dp=None, accumulate=False):
        self.mesh, self.op, self.B, self.pos = mesh, op, B, mesh.phi_pos
        self.x, self.y, self.target, self.dy, self.gx, self.dp = x, y, target, dy, gx, dp
        self.accumulate = accumulate
        N, dev = mesh.N, mesh.device
        self.n_m = mesh.n_cr * mesh.X_np
        self.phi = mesh.p_phase[self.n_m:]
        self.nphi = torch.empty_like(self.phi)
        self.ws_n = _lib.phase_screen_workspace_size(N, B)
        self.ws = torch.empty(max(self.ws_n, 1), dtype=torch.uint8, device=dev)
        f = lambda: torch.empty((B, N), dtype=torch.complex64, device=dev)  # noqa: E731
        self.sub = {}
        if op == "fwd" and self.pos == "in":
            self.xs = f()
            self.sub["u_in"] = self.xs
        elif op == "grad" and self.pos == "in":
            self.xs, self.gxs = f(), f()
            self.sub.update(u_in=self.xs, dJdu_in=self.gxs)
        elif op == "grad":
            self.ts, self.gbuf, self.ys = f(), f(), f()
            self.sub["target" if target is not None else "dJdu_out"] = self.ts
        elif op == "backdiff" and self.pos == "out":
            self.ys, self.gbuf = f(), f()
            self.sub.update(u_out=self.ys, dJdu_out=self.gbuf, target=None)
        elif op == "backdiff":
            self.xs, self.gxs = f(), f()
            self.sub.update(u_in=self.xs, dJdu_in=self.gxs)

    def _scr(self, op, phi, u, du=None):
        dphi = self.dp[self.n_m:] if op == _lib.OP_BACKDIFF else None
        _lib.phase_screen(op, self.mesh.N, self.B, phi, dphi, u, du, self.mesh.N, workspace=self.ws,
                          ws_bytes=self.ws_n)

    def _grad_out(self):
        """dJ/dy into gbuf: 2 (y - target) or the given dJ/dy."""
        if self.target is not None:
            torch.sub(self.y, self.target, out=self.gbuf)
            self.gbuf.mul_(2)
        else:
            self.gbuf.copy_(self.dy)

    def pre(self):
        if self.B == 0:
            return
        if self.op != "fwd" and not self.accumulate:
            self.dp[self.n_m:].zero_()  # the screen's adjoint adds into its part
        if self.op in ("fwd", "grad") and self.pos == "in":
            self.xs.copy_(self.x)
            self._scr(_lib.OP_FWDPROP, self.phi, self.xs)
        elif self.op == "grad":
            torch.neg(self.phi, out=self.nphi)
            self.ts.copy_(self.target if self.target is not None else self.dy)
            self._scr(_lib.OP_FWDPROP, self.nphi, self.ts)  # S^H target (or S^H dJ/dy)
        elif self.op == "backdiff" and self.pos == "out":
            self.ys.copy_(self.y)
            self._grad_out()
            self._scr(_lib.OP_BACKDIFF, self.phi, self.ys, self.gbuf)

    def post(self):
        if self.B == 0:
            return
        if self.op == "fwd":
            if self.pos == "out":
                self._scr(_lib.OP_FWDPROP, self.phi, self.y)
        elif self.pos == "in":  # grad and backdiff: back through the input screen
            self._scr(_lib.OP_BACKDIFF, self.phi, self.xs, self.gxs)
            if self.op == "backdiff" and self.x is not None:
                self.x.copy_(self.xs)
            if self.gx is not None:
                self.gx.copy_(self.gxs)
        elif self.op == "grad":
            self._scr(_lib.OP_FWDPROP, self.phi, self.y)  # y = S (M x)
            self._grad_out()
            self.ys.copy_(self.y)
            self._scr(_lib.OP_BACKDIFF, self.phi, self.ys, self.gbuf)


def tables_for(mesh: MeshNetworkGPU, tile=None) -> torch.Tensor:
    """A device buffer for a mesh's resident row tables (kept across group launches)."""
    lib = _lib.load()
    t = _lib.Tile(*((list(tile or []) + [0, 0, 0, 0])[:4]))
    n = int(lib.meshprop_group_tables_size(_lib.X_CODE[mesh.X], mesh.N, mesh.L, int(mesh.fft), ctypes.byref(t)))
    if n == 0:
        raise _lib.MeshpropError("tables_size: unsupported mesh")
    return torch.empty(n + 256, dtype=torch.uint8, device=mesh.device)


def _aligned_ptr(t: torch.Tensor) -> int:
    p = t.data_ptr()
    return (p + 255) & ~255


def _entry(job: GroupJob) -> _lib.GroupEntry:
    m = job.mesh
    e = _lib.GroupEntry()
    e.op = OPS[job.op]
    e.X = _lib.X_CODE[m.X]
    e.N, e.L, e.B, e.fft = m.N, m.L, job.B, int(m.fft)
    g0, g1 = m._geom()
    e.geom0 = g0.data_ptr() if (g0 is not None and g0.numel()) else None
    e.geom1 = g1.data_ptr() if (g1 is not None and g1.numel()) else None
    e.p = m.p_phase.data_ptr() if m.p_phase.numel() else None
    e.ldp = m.N // 2 * m.X_np
    e.s = m.p_splitter.data_ptr() if m.p_splitter.numel() else None
    e.lds = m.N // 2 * m.X_ns
    e.dp = job.dp.data_ptr() if job.dp is not None else None
    if job.op == "backdiff":
        e.u_in = job.x.data_ptr() if job.x is not None else None
    else:
        e.u_in = job.x.data_ptr()
    e.u_out = job.y.data_ptr()
    e.target = job.target.data_ptr() if job.target is not None else None
    e.dJdu_out = job.dy.data_ptr() if job.dy is not None else None
    e.dJdu_in = job.gx.data_ptr() if job.gx is not None else None
    e.ldu = m.N
    w, v, waves, chunk = (list(job.tile or []) + [0, 0, 0, 0])[:4]
    e.tile = _lib.Tile(int(w), int(v), int(waves), int(chunk))
    e.tables = _aligned_ptr(job.tables) if job.tables is not None else None
    e.build_tables = int(bool(job.build_tables))
    e.flags = 0 if job.accumulate else _lib.GROUP_DP_ASSIGN
    if job.screen is not None:
        for name, t in job.screen.sub.items():
            setattr(e, name, t.data_ptr() if t is not None else None)
    return e


class MeshGroup:
    """Plan, workspace and launches of a grouped propagation of ``jobs``.

    ``run()`` launches eagerly on the current stream; ``capture()`` records the launches as a
    hipGraph (through torch.cuda.CUDAGraph) and ``run()`` then replays it.  ``replan(jobs)``
    uploads a new plan into the same workspace (growing it if needed) and keeps the graph when
    the launch signature is unchanged.
    """

    def __init__(self, jobs: List[GroupJob], workspace: Optional[torch.Tensor] = None):
        self._lib = _lib.load()
        self._plan = ctypes.c_void_p()
        self._graphs = {}  # launch signature -> captured graph
        self._ws = workspace
        self.jobs: List[GroupJob] = []
        self.replan(jobs)

    # -- plan ------------------------------------------------------------------------------
    def replan(self, jobs: List[GroupJob]):
        self.jobs = list(jobs)
        ents = (_lib.GroupEntry * max(1, len(self.jobs)))(*[_entry(j) for j in self.jobs])
        dev = self.jobs[0].mesh.device if self.jobs else None
        self.replan_entries(ents, len(self.jobs), dev)

    def replan_entries(self, ents, n: int, device=None):
        """Plan from a ctypes array of meshprop_group_entry (the fast path of the server)."""
        plan = ctypes.c_void_p()
        _lib.check(self._lib.meshprop_group_create(ents, n, ctypes.byref(plan)), "group_create")
        self._free_plan()
        self._plan = plan
        self._entries = ents
        need = int(self._lib.meshprop_group_workspace_size(self._plan))
        if self._ws is None or self._ws.numel() < need + 256:
            dev = device if device is not None else torch.device("cuda", torch.cuda.current_device())
            self._ws = torch.empty(need + 256, dtype=torch.uint8, device=dev)
            self._graphs = {}  # captured graphs hold the old workspace address
        self.ws_bytes = need
        self.signature = int(self._lib.meshprop_group_signature(self._plan))
        self.upload()

    @property
    def workspace_bytes(self) -> int:
        return self._ws.numel() - 256

    def upload(self):
        s = torch.cuda.current_stream().cuda_stream
        _lib.check(self._lib.meshprop_group_upload(self._plan, _aligned_ptr(self._ws), self.ws_bytes,
                                                   ctypes.c_void_p(s)), "group_upload")

    def info(self, i: int):
        """(tile (w, v, waves, chunk), number of work items) of job i."""
        t = _lib.Tile()
        n = ctypes.c_int()
        _lib.check(self._lib.meshprop_group_entry_info(self._plan, i, ctypes.byref(t), ctypes.byref(n)), "info")
        return (t.w, t.v, t.waves, t.chunk), n.value

    # -- launches --------------------------------------------------------------------------
    def launch(self):
        s = torch.cuda.current_stream().cuda_stream
        screens = [j.screen for j in self.jobs if j.screen is not None]
        for sp in screens:
            sp.pre()
        _lib.check(self._lib.meshprop_group_launch(self._plan, _aligned_ptr(self._ws), self.ws_bytes,
                                                   ctypes.c_void_p(s)), "group_launch")
        for sp in screens:
            sp.post()

    def capture(self, warm: bool = True):
        """Record prologue, persistent kernel and reduction as one replayable graph (kept per
        launch signature; later plans uploaded into this workspace replay it)."""
        if warm:
            self.launch()  # loads the code objects before capture
            torch.cuda.synchronize()
            self.upload()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.launch()
        torch.cuda.synchronize()
        self._graphs[self.signature] = g

    @property
    def graph(self):
        return self._graphs.get(self.signature)

    def run(self):
        g = self._graphs.get(self.signature)
        if g is not None:
            g.replay()
        else:
            self.launch()

    def _free_plan(self):
        if self._plan:
            self._lib.meshprop_group_destroy(self._plan)
            self._plan = ctypes.c_void_p()

    def close(self):
        self._graphs = {}
        self._free_plan()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PlainRunner:
    """The per-mesh path: each job as its own plain launch sequence (prologue, propagation,
    reduction) through meshprop_run, one mesh after another on one stream, with workspaces
    allocated once up front."""

    def __init__(self, jobs: List[GroupJob]):
        self.jobs = list(jobs)
        self.tiles = [j.tile for j in self.jobs]
        self.ws = []
        for j in self.jobs:
            m = j.mesh
            ops = {"fwd": [_lib.OP_FWDPROP], "grad": [_lib.OP_FWDPROP, _lib.OP_BACKDIFF],
                   "backdiff": [_lib.OP_BACKDIFF]}[j.op]
            n = max(_lib.workspace_size(op, m.X, m.N, m.L, j.B, m.fft, j.tile) for op in ops)
            # meshprop_run's backdiff always writes the reconstructed input: scratch for jobs without x
            rec = j.x if j.op == "backdiff" and j.x is not None else (
                torch.empty_like(j.y) if j.op != "fwd" else None)
            self.ws.append((torch.empty(max(n, 1), dtype=torch.uint8, device=m.device), n, rec))

    def run(self):
        for j, tile, (ws, nb, rec) in zip(self.jobs, self.tiles, self.ws):
            m = j.mesh
            g0, g1 = m._geom()
            ldp, lds = m.N // 2 * m.X_np, m.N // 2 * m.X_ns
            s = m.p_splitter if m.p_splitter.numel() else None
            if j.B == 0:
                continue
            # the job's buffers, or the screen plan's in their place
            f = dict(u_in=j.x, u_out=j.y, target=j.target, dJdu_out=j.dy, dJdu_in=j.gx)
            if j.screen is not None:
                f.update(j.screen.sub)
                if not j.accumulate and j.op != "fwd":
                    j.dp.zero_()
                j.screen.pre()
            if j.op in ("fwd", "grad"):
                _lib.run(_lib.OP_FWDPROP, m.X, m.N, m.L, j.B, m.fft, g0, g1, m.p_phase, None, ldp, s, lds,
                         f["u_in"], None, j.y, None, m.N, tile=tile, workspace=ws, ws_bytes=nb)
            if j.op in ("grad", "backdiff"):
                if not j.accumulate and j.screen is None:
                    j.dp.zero_()
                u_rec = f["u_in"] if (j.op == "backdiff" and f["u_in"] is not None) else rec
                if j.op == "grad":
                    u_rec = rec
                _lib.run(_lib.OP_BACKDIFF, m.X, m.N, m.L, j.B, m.fft, g0, g1, m.p_phase, j.dp, ldp, s, lds,
                         f["u_out"], f["dJdu_out"], u_rec, f["dJdu_in"], m.N, target=f["target"], tile=tile,
                         workspace=ws, ws_bytes=nb)
            if j.screen is not None:
                j.screen.post()


def run_plain(jobs: List[GroupJob]):
    PlainRunner(jobs).run()
