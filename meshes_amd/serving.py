"""Serving: continuous batching of forward and gradient requests over device-resident meshes.

``MeshServer`` holds a set of meshes whose phases and row tables stay in HBM (the tables are
built once, at registration, by a grouped prologue).  Requests (mesh id, 'fwd' or 'grad', a
(B, N) input and, for gradients, a target) arrive over time.  The request loop keeps up to
``slots`` grouped launches in flight, each on its own stream with its own workspace:
whenever a slot is free it takes the waiting requests in arrival order, as many as the HBM
budget allows (the footprint of a request is its field buffers plus its dp partials), plans
them as one grouped launch, uploads the plan and replays the slot's captured graph.  Requests
that arrive while batches run wait only for a free slot and memory, and consecutive batches
overlap on the device (the workgroups of the next persistent kernel start as those of the
previous one drain).  Completion is observed by polling each slot's event.

A gradient request returns dp (and y): the gradient of J = sum |U x - target|^2 with respect
to the mesh phases, exactly as a grouped 'grad' job computes it; a forward request returns
y = U x.  ``serve_sequential`` is the baseline: the same stream of requests served one at a
time with the plain per-call launches (gpu/mesh.py:158-231 call pattern), waiting for each.
"""
from __future__ import annotations

import collections
import ctypes
import time
from dataclasses import dataclass
from typing import List, Optional

import numpy as np
import torch

from . import _lib
from .batched import GroupJob, MeshGroup, ScreenPlan, _entry, tables_for
from .mesh import MeshNetworkGPU


@dataclass
class Request:
    rid: int
    mesh_id: int
    op: str                       # 'fwd' or 'grad'
    x: torch.Tensor               # (B, N) complex64 on the device (float32 for orth meshes)
    target: Optional[torch.Tensor] = None
    t_arrive: float = 0.0         # seconds after the stream starts
    y: Optional[torch.Tensor] = None
    dp: Optional[torch.Tensor] = None
    t_done: float = -1.0
    footprint: int = 0
    screen: Optional[ScreenPlan] = None  # launches of the mesh's phase screen, if it has one

    @property
    def B(self):
        return int(self.x.shape[0])

    @property
    def latency(self):
        return self.t_done - self.t_arrive


def _lpv(N, W):
    p = 1
    while p < (N + W - 1) // W:
        p *= 2
    return p


def _width(N):
    return 2 if N <= 128 else 4 if N <= 256 else 8 if N <= 512 else 16


class _Slot:
    def __init__(self, device, capacity):
        self.stream = torch.cuda.Stream(device=device)
        self.workspace = torch.empty(capacity + 256, dtype=torch.uint8, device=device)
        self.group: Optional[MeshGroup] = None
        self.event = torch.cuda.Event()
        self.reqs: List[Request] = []
        self.bytes = 0

    @property
    def busy(self):
        return bool(self.reqs)


class MeshServer:
    """Continuous-batching server over device-resident meshes.

    :param meshes: the resident meshes (complex, unitary crossings for gradient requests).
    :param hbm_budget: bytes of HBM the in-flight requests may hold (field buffers, dp partials
        and per-slot workspaces); a request waits until it fits.
    :param slots: grouped launches in flight at once (one stream and workspace each).
    :param max_batch: most requests in one grouped launch.
    """

    def __init__(self, meshes: List[MeshNetworkGPU], hbm_budget: int = 8 << 30, slots: int = 3,
                 max_batch: int = 256, graphs: bool = True, direct: bool = True):
        _lib.load()
        self.meshes = list(meshes)
        self.device = self.meshes[0].device
        self.budget = int(hbm_budget)
        self.max_batch = int(max_batch)
        self.ncu = torch.cuda.get_device_properties(self.device).multi_processor_count
        self.slot_capacity = self.budget // max(1, slots)
        self.slots = [_Slot(self.device, self.slot_capacity) for _ in range(slots)]
        self.inflight_bytes = 0
        self.graphs = graphs
        self.direct = direct
        self.direct_launches = 0
        # Row tables of every mesh, built once and kept resident.
        self.tables = [tables_for(m) for m in self.meshes]
        build = [GroupJob(m, "fwd", x=torch.empty((0, m.N), dtype=m.dtype, device=self.device),
                          tables=t, build_tables=True) for m, t in zip(self.meshes, self.tables)]
        g = MeshGroup(build)
        g.run()
        torch.cuda.synchronize()
        g.close()
        # Per-mesh entry templates (everything but the request's op, batch and buffers).
        self._tmpl = []
        for m, t in zip(self.meshes, self.tables):
            e = _entry(GroupJob(m, "fwd", x=torch.empty((0, m.N), dtype=m.dtype, device=self.device),
                                tables=t, build_tables=False))
            e.flags = _lib.GROUP_DP_ASSIGN
            self._tmpl.append(e)
        self._esize = ctypes.sizeof(_lib.GroupEntry)
        if graphs:
            self._capture_slots()

    # -- footprint --------------------------------------------------------------------------
    def footprint(self, mesh_id: int, op: str, B: int) -> int:
        return request_footprint(self.meshes[mesh_id], op, B, self.ncu)

    # -- graphs -----------------------------------------------------------------------------
    @staticmethod
    def _family(m) -> int:
        """Persistent kernel family of a mesh's grouped items: 0 for N <= 512, 1 for wider NN
        meshes of MZI / SYM crossings, 2 for the other wide meshes."""
        if m.N <= 512:
            return 0
        return 1 if (not m.fft and m.X in ("mzi", "sym")) else 2

    def _capture_slots(self):
        """Capture each slot's launch once per signature a batch can have: forward-only or
        with gradients, with items of any non-empty set of the kernel families present."""
        rep = {}
        for i, m in enumerate(self.meshes):
            rep.setdefault(self._family(m), i)
        fams = sorted(rep)
        combos = []
        for op in ("fwd", "grad"):
            for mask in range(1, 1 << len(fams)):
                combos.append([(rep[f], op) for k, f in enumerate(fams) if mask >> k & 1])
        for s in self.slots:
            with torch.cuda.stream(s.stream):
                for combo in combos:
                    reqs = []
                    for mid, op in combo:
                        N = self.meshes[mid].N
                        x = torch.zeros((1, N), dtype=self.meshes[mid].dtype, device=self.device)
                        reqs.append(Request(-1, mid, op, x, torch.zeros_like(x) if op == "grad" else None))
                    self._plan(s, reqs)
                    s.group.capture()
            torch.cuda.synchronize()
            s.reqs, s.bytes = [], 0

    # -- request loop -----------------------------------------------------------------------
    def _plan(self, slot: _Slot, batch: List[Request]):
        """Allocate the batch's outputs and upload its grouped plan into the slot's workspace
        (on the slot's stream)."""
        n = len(batch)
        ents = (_lib.GroupEntry * n)()
        base = ctypes.addressof(ents)
        with torch.cuda.stream(slot.stream):
            for k, r in enumerate(batch):
                m = self.meshes[r.mesh_id]
                ctypes.memmove(base + k * self._esize, ctypes.addressof(self._tmpl[r.mesh_id]), self._esize)
                e = ents[k]
                e.B = r.B
                e.u_in = r.x.data_ptr()
                r.y = torch.empty((r.B, m.N), dtype=m.dtype, device=self.device)
                e.u_out = r.y.data_ptr()
                if r.op == "grad":
                    r.dp = torch.empty(m.n_phase, dtype=torch.float32, device=self.device)
                    e.op = _lib.OP_FWDGRAD
                    e.dp = r.dp.data_ptr()
                    e.target = r.target.data_ptr()
                else:
                    e.op = _lib.OP_FWDPROP
                # a phase screen runs as screen launches around the grouped launch
                r.screen = ScreenPlan(m, r.op, r.B, x=r.x, y=r.y, target=r.target, dp=r.dp) if m.is_phase else None
                if r.screen is not None:
                    for name, t in r.screen.sub.items():
                        setattr(e, name, t.data_ptr() if t is not None else None)
            if slot.group is None:
                slot.group = MeshGroup([], workspace=slot.workspace)
            slot.group.replan_entries(ents, n, self.device)
            if slot.group.workspace_bytes != self.slot_capacity:
                raise RuntimeError("a batch outgrew its slot workspace")
        slot.reqs = batch

    def _launch_direct(self, slot: _Slot, r: Request) -> bool:
        """A lone request (the other slots idle) as one plain meshprop_run call on the slot's
        stream, with the slot's workspace: nothing to batch with, so no plan upload and no
        grouped launch.  False when its workspace would not fit the slot's."""
        m = self.meshes[r.mesh_id]
        op = _lib.OP_FWDGRAD if r.op == "grad" else _lib.OP_FWDPROP
        need = _lib.workspace_size(op, m.X, m.N, m.L, r.B, m.fft)
        if m.is_phase or need > slot.workspace.numel():
            return False
        g0, g1 = m._geom()
        ldp, lds = m.N // 2 * m.X_np, m.N // 2 * m.X_ns
        sp = m.p_splitter if m.p_splitter.numel() else None
        with torch.cuda.stream(slot.stream):
            r.y = torch.empty((r.B, m.N), dtype=m.dtype, device=self.device)
            r.screen = None
            if r.op == "grad":
                r.dp = torch.empty(m.n_phase, dtype=torch.float32, device=self.device)
                _lib.run(op, m.X, m.N, m.L, r.B, m.fft, g0, g1, m.p_phase, r.dp, ldp, sp, lds, r.x, None, r.y, None,
                         m.N, target=r.target, workspace=slot.workspace, ws_bytes=need, flags=_lib.DP_ASSIGN)
            else:
                _lib.run(op, m.X, m.N, m.L, r.B, m.fft, g0, g1, m.p_phase, None, ldp, sp, lds, r.x, None, r.y, None,
                         m.N, workspace=slot.workspace, ws_bytes=need)
            slot.event.record(slot.stream)
        slot.reqs = [r]
        self.direct_launches += 1
        return True

    def _launch(self, slot: _Slot, batch: List[Request]):
        self._plan(slot, batch)
        screens = [r.screen for r in batch if r.screen is not None]
        with torch.cuda.stream(slot.stream):
            for sp in screens:
                sp.pre()
            slot.group.run()
            for sp in screens:
                sp.post()
            slot.event.record(slot.stream)

    def serve(self, requests: List[Request], clock=time.perf_counter):
        """Run the request stream in real time (arrival times relative to the start); returns
        the requests with t_done set and a statistics dict."""
        reqs = sorted(requests, key=lambda r: r.t_arrive)
        for r in reqs:
            r.x = r.x.resolve_conj().resolve_neg()
            if r.target is not None:
                r.target = r.target.resolve_conj().resolve_neg()
            r.footprint = self.footprint(r.mesh_id, r.op, r.B)
            if r.footprint > self.slot_capacity:
                raise ValueError(f"request {r.rid} needs {r.footprint} B, above one slot's {self.slot_capacity} B")
        torch.cuda.synchronize()
        self.direct_launches = 0
        pending = collections.deque()
        n, i, done, batches, peak = len(reqs), 0, 0, 0, 0
        sizes = []
        t0 = clock()
        while done < n:
            now = clock() - t0
            while i < n and reqs[i].t_arrive <= now:
                pending.append(reqs[i])
                i += 1
            for s in self.slots:  # retire finished batches
                if s.busy and s.event.query():
                    tdone = clock() - t0
                    for r in s.reqs:
                        r.t_done = tdone
                    done += len(s.reqs)
                    self.inflight_bytes -= s.bytes
                    s.reqs, s.bytes = [], 0
            for s in self.slots:  # admit into free slots
                if s.busy or not pending:
                    continue
                batch, nbytes = [], 0
                while pending and len(batch) < self.max_batch:
                    f = pending[0].footprint
                    if nbytes + f > self.slot_capacity or self.inflight_bytes + nbytes + f > self.budget:
                        break
                    batch.append(pending.popleft())
                    nbytes += f
                if not batch:
                    break
                s.bytes = nbytes
                self.inflight_bytes += nbytes
                peak = max(peak, self.inflight_bytes)
                lone = len(batch) == 1 and not any(o.busy for o in self.slots if o is not s)
                if not (self.direct and lone and self._launch_direct(s, batch[0])):
                    self._launch(s, batch)
                batches += 1
                sizes.append(len(batch))
        torch.cuda.synchronize()
        return reqs, _stats(reqs, self.meshes, batches=batches, direct_launches=self.direct_launches,
                            mean_batch=float(np.mean(sizes)) if sizes else 0.0,
                            peak_inflight_bytes=peak)

    # -- baseline ---------------------------------------------------------------------------
    def serve_sequential(self, requests: List[Request], clock=time.perf_counter):
        """The same stream served one request at a time with plain launches (prologue,
        propagation, reduction per call), each waited for before the next starts."""
        reqs = sorted(requests, key=lambda r: r.t_arrive)
        ws = torch.empty(1 << 20, dtype=torch.uint8, device=self.device)
        torch.cuda.synchronize()
        t0 = clock()
        for r in reqs:
            while clock() - t0 < r.t_arrive:
                pass
            m = self.meshes[r.mesh_id]
            if m.is_phase:  # the mesh's own methods put the screen around its launches
                if r.op == "grad":
                    _, r.dp, uo = m.grad_target_l2(r.x.T, r.target.T)
                    r.y = uo.T
                else:
                    r.y = m.dot(r.x.T).T
                torch.cuda.current_stream().synchronize()
                r.t_done = clock() - t0
                continue
            g0, g1 = m._geom()
            ldp, lds = m.N // 2 * m.X_np, m.N // 2 * m.X_ns
            s = m.p_splitter if m.p_splitter.numel() else None  # as the served entries carry them
            ops = [_lib.OP_FWDPROP] + ([_lib.OP_BACKDIFF] if r.op == "grad" else [])
            need = max(_lib.workspace_size(op, m.X, m.N, m.L, r.B, m.fft) for op in ops)
            if ws.numel() < need:
                ws = torch.empty(need, dtype=torch.uint8, device=self.device)
            r.y = torch.empty((r.B, m.N), dtype=m.dtype, device=self.device)
            _lib.run(_lib.OP_FWDPROP, m.X, m.N, m.L, r.B, m.fft, g0, g1, m.p_phase, None, ldp, s, lds, r.x, None,
                     r.y, None, m.N, workspace=ws, ws_bytes=need)
            if r.op == "grad":
                r.dp = torch.zeros(m.n_phase, dtype=torch.float32, device=self.device)
                rec = torch.empty_like(r.y)
                _lib.run(_lib.OP_BACKDIFF, m.X, m.N, m.L, r.B, m.fft, g0, g1, m.p_phase, r.dp, ldp, s, lds, r.y,
                         None, rec, None, m.N, target=r.target, workspace=ws, ws_bytes=need)
            torch.cuda.current_stream().synchronize()
            r.t_done = clock() - t0
        return reqs, _stats(reqs, self.meshes)


def request_footprint(m: MeshNetworkGPU, op: str, B: int, ncu: int) -> int:
    """HBM bytes a request holds while in flight: its field buffers (x, y and the target)
    and, for gradients, its dp and dp partials (one slab per partial slot of the grouped
    plan: min(vector tiles, 2 x CUs) slabs of L x N/2 x 2 floats), plus its share of the
    plan (descriptor, queue entries, alignment)."""
    field_bytes = B * m.N * m.dtype.itemsize
    W = _width(m.N)
    lpv = _lpv(m.N, W)
    waves = 2 if W == 16 else 4
    per_tile = waves * (64 // lpv)
    tiles = (B + per_tile - 1) // per_tile
    plan = 2048
    # a phase screen's buffers: up to three fields, its dphi partials and -phi
    screen = (3 * field_bytes + ((B + 63) // 64) * m.N * 4 + m.N * 4 + 256) if m.is_phase else 0
    if op == "fwd":
        return 2 * field_bytes + plan + 8 * tiles + screen
    slabs = min(tiles, 2 * ncu)
    partial = slabs * m.L * (lpv * W // 2) * 2 * 4
    return 3 * field_bytes + m.n_phase * 4 + partial + plan + 8 * slabs + screen


def _stats(reqs, meshes, **extra):
    lat = np.array([r.latency for r in reqs]) * 1e3
    span = max(r.t_done for r in reqs) - min(r.t_arrive for r in reqs)
    upd = 0
    for r in reqs:
        m = meshes[r.mesh_id]
        act = r.B * (m.L * (m.N // 2) if m.fft else int(m.lens.sum().item()))
        upd += act
    out = {"requests": len(reqs), "span_s": span, "throughput_req_s": len(reqs) / span,
           "updates_per_s": upd / span, "p50_ms": float(np.percentile(lat, 50)),
           "p99_ms": float(np.percentile(lat, 99)), "mean_ms": float(lat.mean())}
    out.update(extra)
    return out


def poisson_requests(meshes: List[MeshNetworkGPU], rate: float, n: int, seed: int = 0, grad_frac: float = 0.3,
                     b_median: float = 64.0, b_sigma: float = 1.5, b_max: int = 8192, device=None):
    """n requests with exponential inter-arrival times at ``rate`` per second; each picks a
    resident mesh uniformly, is a gradient job with probability grad_frac, and carries a
    lognormal batch (median b_median, sigma b_sigma, clipped to [1, b_max]) of unit-norm
    CN(0, 1) vectors (and a unit-norm target for gradients), already on the device."""
    rng = np.random.RandomState(seed)
    dev = device or meshes[0].device
    t = np.cumsum(rng.exponential(1.0 / rate, size=n))
    t -= t[0]
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    out = []
    for k in range(n):
        mid = int(rng.randint(len(meshes)))
        m = meshes[mid]
        B = int(np.clip(np.round(b_median * np.exp(b_sigma * rng.standard_normal())), 1, b_max))
        op = "grad" if rng.uniform() < grad_frac else "fwd"

        def unit():
            z = torch.randn((B, m.N), dtype=m.dtype, device=dev, generator=gen)
            return z / torch.linalg.norm(z, dim=1, keepdim=True)
        out.append(Request(k, mid, op, unit(), unit() if op == "grad" else None, float(t[k])))
    return out


def resident_meshes(n: int = 256, n_min: int = 16, n_max: int = 1024, seed: int = 0) -> List[MeshNetworkGPU]:
    """n Clements MZI meshes (depth N) with N log-uniform in [n_min, n_max] (rounded to even),
    phases U[0, 2 pi)."""
    from . import geometry
    rng = np.random.RandomState(seed)
    out = []
    for _ in range(n):
        N = int(2 * round(np.exp(rng.uniform(np.log(n_min), np.log(n_max))) / 2))
        N = min(max(N, n_min), n_max)
        lens, shifts = geometry.clements(N)
        p = geometry.uniform_phases(rng, N, N)
        out.append(MeshNetworkGPU(N, N, lens, shifts, p, np.zeros(0), "mzi", is_phase=False))
    return out
