"""Measured tile table of the plain path (meshprop_run called without a tile).

The reference picks the warps per block of each kernel from measured tables keyed by
(kernel, crossing type, square or fat batch, N bucket) (gpu/__init__.py:26-35, written by the
sweeps of gpu/test.py:163-250): "square" is an N x N x N workload, "fat" N x N x 4096, and a call
counts as fat when B > 2N (gpu/mesh.py:161, 171, 229), with N rounded up to the next tabulated
size (gpu/mesh.py:84).  This sweep does the same for the tiles of this library -- W waveguides
per lane, V vectors per lane and waves per workgroup -- over the four operations (fwdprop,
fwddiff, backdiff and the fused fwd+grad) and the four crossing types, timing whole library calls
(prologue, propagation and dp reduction) with HIP events on the calling stream.  The winners are
written as the C++ initializer list csrc/tile_table.inc that resolve_tile() consults; a Markdown
log records every case with the heuristic's time beside the winner's.

    python -m meshes_amd.tune_tiles --inc meshes_amd/csrc/tile_table.inc --log tile_sweep.md

Runs on one GPU in a few minutes and prints one line per case.
"""
from __future__ import annotations

import argparse
import datetime
import sys
import time

import numpy as np
import torch

from . import _lib as Lb
from . import geometry

OPS = {"fwdprop": Lb.OP_FWDPROP, "fwddiff": Lb.OP_FWDDIFF, "backdiff": Lb.OP_BACKDIFF, "fwdgrad": Lb.OP_FWDGRAD}
XS = ["mzi", "sym", "orth", "mzi_generic"]
N_BUCKETS = [8, 16, 32, 64, 128, 256, 512, 1024]
FAT_B = 4096  # the reference's fat workload (gpu/test.py:238)
MARGIN = 0.03  # least gain over the current choice that changes the table


def _pow2ceil(x: int) -> int:
    p = 1
    while p < x:
        p <<= 1
    return p


def candidates(op: int, N: int):
    """Tiles the library compiles for this op and width: the narrowest width that holds a
    vector in one wave and the next wider one, every compiled V, and 1-8 waves per workgroup
    (the plan clamps what an adjoint cannot run)."""
    wmin = next(w for w in (2, 4, 8, 16) if w * 64 >= N)
    out = []
    for W in (2, 4, 8, 16):
        if W < wmin or W > 2 * wmin:
            continue
        if _pow2ceil(-(-N // W)) > 64:
            continue
        for V in (1, 2, 4, 8):
            ok = (W * V <= 32 and V <= 8) if op == Lb.OP_FWDPROP else (W * V <= 16 and V <= 4)
            if not ok:
                continue
            for waves in (1, 2, 4, 8):
                out.append((W, V, waves, 0))
    return out


class Case:
    """Inputs of one (op, X, N, B) timing case on the current device."""

    def __init__(self, op: int, X: str, N: int, L: int, B: int, seed: int = 0):
        rng = np.random.RandomState(seed)
        self.op, self.X, self.N, self.L, self.B = op, X, N, L, B
        lens, shifts = geometry.clements(N)
        # depth L: the Clements layer pair repeated (the first L layers when L <= N)
        lens = [int(lens[l % 2]) for l in range(L)]
        shifts = [int(shifts[l % 2]) for l in range(L)]
        self.n_active = int(sum(lens))
        dev = torch.device("cuda", 0)
        self.lens = torch.as_tensor(np.asarray(lens, np.int32)).to(dev)
        self.shifts = torch.as_tensor(np.asarray(shifts, np.int32)).to(dev)
        npar, nspl = Lb.N_PHASE[X], Lb.N_SPLIT[X]
        self.ldp, self.lds = (N // 2) * npar, (N // 2) * nspl
        self.p = torch.as_tensor(rng.uniform(0, 2 * np.pi, (L, N // 2, npar)).astype(np.float32)).to(dev)
        s = np.zeros((L, N // 2, nspl), np.float32)
        if X == "mzi_generic":
            s[..., :2] = 0.02 * rng.standard_normal((L, N // 2, 2))
            s[..., 10:] = np.float32(np.log(10 ** (-0.1 / 20)))
        elif nspl:
            s[...] = 0.02 * rng.standard_normal(s.shape)
        self.s = torch.as_tensor(s).to(dev) if nspl else None
        real = X == "orth"
        dt = torch.float32 if real else torch.complex64

        def field():
            a = rng.standard_normal((B, N))
            if not real:
                a = a + 1j * rng.standard_normal((B, N))
            # the kernels read complex64 (float32 for real fields): cast after the scaling
            return torch.as_tensor((a / np.sqrt(N)).astype(np.float32 if real else np.complex64)).to(dev)

        self.u, self.du, self.y = field(), field(), field()
        self.out, self.dout = torch.empty_like(self.u), torch.empty_like(self.u)
        self.dp = torch.zeros((L, N // 2, npar), dtype=torch.float32, device=dev)
        self.flags = Lb.SCALED_UNITARY if (X == "mzi_generic" and op == Lb.OP_FWDGRAD) else 0
        self.extra_flags = 0  # e.g. GENERIC_KERNELS, to time the general kernels
        self.dt = dt

    def run(self, tile, ws, nb):
        op, X, N, L, B = self.op, self.X, self.N, self.L, self.B
        g = (self.lens, self.shifts)
        if op == Lb.OP_FWDPROP:
            Lb.run(op, X, N, L, B, False, *g, self.p, None, self.ldp, self.s, self.lds, self.u, None, self.out, None,
                   N, tile=tile, workspace=ws, ws_bytes=nb, flags=self.extra_flags)
        elif op == Lb.OP_FWDDIFF:
            Lb.run(op, X, N, L, B, False, *g, self.p, self.dp, self.ldp, self.s, self.lds, self.u, self.du, self.out,
                   self.dout, N, tile=tile, workspace=ws, ws_bytes=nb, flags=self.extra_flags)
        elif op == Lb.OP_BACKDIFF:
            Lb.run(op, X, N, L, B, False, *g, self.p, self.dp, self.ldp, self.s, self.lds, self.u, None, self.out,
                   None, N, target=self.y, tile=tile, workspace=ws, ws_bytes=nb, flags=self.extra_flags)
        else:
            Lb.run(op, X, N, L, B, False, *g, self.p, self.dp, self.ldp, self.s, self.lds, self.u, None, self.out,
                   None, N, target=self.y, tile=tile, workspace=ws, ws_bytes=nb,
                   flags=self.flags | self.extra_flags | Lb.DP_ASSIGN)

    def time(self, tile, reps: int = 5, span_ms: float = 0.5) -> float:
        """Median ms of one call with this tile (inf when the library rejects the tile): each of
        ``reps`` samples times enough back-to-back calls to span about ``span_ms``."""
        try:
            nb = Lb.workspace_size(self.op, self.X, self.N, self.L, self.B, False, tile)
            ws = torch.empty(max(nb, 256), dtype=torch.uint8, device="cuda")
            self.run(tile, ws, nb)
        except RuntimeError:
            return float("inf")
        self.run(tile, ws, nb)
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        self.run(tile, ws, nb)
        b.record()
        b.synchronize()
        k = max(1, min(64, int(span_ms / max(a.elapsed_time(b), 1e-3))))
        ts = []
        for _ in range(reps):
            a.record()
            for _ in range(k):
                self.run(tile, ws, nb)
            b.record()
            b.synchronize()
            ts.append(a.elapsed_time(b) / k)
        return float(np.median(ts))


def _resolved(op, X, N, B, tile):
    """(W, V, waves) a call runs with this tile (None: the call's default)."""
    try:
        return tuple(Lb.resolve_tile(op, X, N, B, tile)[:3])
    except RuntimeError:  # e.g. fwd+grad of lossy crossings needs the scaled-unitary flag
        return tuple(tile[:3]) if tile else None


def sweep(ops, xs, buckets, log=print, finalists: int = 3, rounds: int = 3):
    """Every distinct tile once, then the fastest few and the current choice again in
    interleaved rounds; the table keeps the best finalist only when its median beats the current
    choice's by MARGIN (one pass over many noisy candidates favours lucky samples)."""
    rows = []
    for opname in ops:
        op = OPS[opname]
        for X in xs:
            for fat in (0, 1):
                for N in buckets:
                    B = FAT_B if fat else N
                    case = Case(op, X, N, N, B)
                    t0 = time.time()
                    base = _resolved(op, X, N, B, None)
                    cands = {}
                    for tile in candidates(op, N):
                        cands.setdefault(_resolved(op, X, N, B, tile), tile)
                    first = {k: case.time(t) for k, t in cands.items()}
                    top = [k for k in sorted(first, key=first.get) if np.isfinite(first[k])][:finalists]
                    runs = {k: [] for k in top}
                    runs["current"] = []
                    for _ in range(rounds):
                        for k in runs:
                            runs[k].append(case.time(None if k == "current" else cands[k]))
                    med = {k: float(np.median(v)) for k, v in runs.items()}
                    t_def = med.pop("current")
                    best = min(med, key=med.get) if med else None
                    t_best = med[best] if best else float("inf")
                    keep = best if (best and t_best * (1 + MARGIN) <= t_def) else base
                    row = dict(op=opname, X=X, fat=fat, N=N, B=B, tile=best, ms=t_best, ms_default=t_def,
                               base=base, keep=keep, secs=time.time() - t0)
                    rows.append(row)
                    log(f"{opname:8s} {X:11s} {'fat' if fat else 'sq ':3s} N={N:5d} B={B:5d} current={base} "
                        f"{t_def:.4f} ms  best={best} {t_best:.4f} ms  table={keep}  ({row['secs']:.1f} s)")
                    del case
                    torch.cuda.empty_cache()
    return rows


def write_inc(rows, path, device):
    with open(path, "w") as f:
        f.write("// Measured tile table of the plain path: {op, X, fat, N bucket, W, V, waves}.\n")
        f.write(f"// Generated by meshes_amd/tune_tiles.py on {device} ({datetime.date.today().isoformat()}):\n")
        f.write("// square N x N x N and fat N x N x 4096 Clements meshes, whole-call time.\n")
        for r in rows:
            if r["keep"] is None:
                continue
            W, V, waves = r["keep"][:3]
            f.write(f"{{{OPS[r['op']]}, {Lb.X_CODE[r['X']]}, {r['fat']}, {r['N']}, {W}, {V}, {waves}}},\n")


def write_log(rows, path, device, cmd):
    with open(path, "w") as f:
        f.write("# Plain-path tile sweep\n\n")
        f.write(f"Device: {device}.  Command: `{cmd}`\n\n")
        f.write("Square = N x N x N (L = B = N), fat = N x N x 4096; Clements geometry; whole meshprop_run "
                "time per call (HIP events over back-to-back calls; every distinct tile once, then the three "
                "fastest and the current choice in three interleaved rounds, medians shown).  `current` is the tile "
                "a call without a tile got before this sweep; `table` is what the sweep wrote (the best tile "
                f"when it is at least {MARGIN:.0%} faster than the current one).\n\n")
        f.write("| op | X | shape | N | current (W,V,waves) | ms | best (W,V,waves) | ms | speedup | table |\n"
                "|---|---|---|---|---|---|---|---|---|---|\n")
        for r in rows:
            t, b, k = r["tile"], r["base"], r["keep"]
            ok = np.isfinite(r["ms"]) and np.isfinite(r["ms_default"]) and r["ms"] > 0
            sp = r["ms_default"] / r["ms"] if ok else float("nan")
            f.write(f"| {r['op']} | {r['X']} | {'fat' if r['fat'] else 'sq'} | {r['N']} | {b} | {r['ms_default']:.4f} | "
                    f"{t} | {r['ms']:.4f} | {sp:.2f} | {k} |\n")


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--inc", required=True, help="C++ initializer list to write")
    ap.add_argument("--log", required=True, help="Markdown log of every case")
    ap.add_argument("--ops", default=",".join(OPS))
    ap.add_argument("--xs", default=",".join(XS))
    ap.add_argument("--buckets", default=",".join(map(str, N_BUCKETS)))
    args = ap.parse_args(argv)
    if not torch.cuda.is_available():
        sys.exit("tune_tiles needs a GPU")
    prop = torch.cuda.get_device_properties(0)
    device = f"{prop.name} ({getattr(prop, 'gcnArchName', '?')}, {prop.multi_processor_count} CUs)"
    rows = sweep(args.ops.split(","), args.xs.split(","), [int(b) for b in args.buckets.split(",")],
                 log=lambda s: print(s, flush=True))
    write_inc(rows, args.inc, device)
    write_log(rows, args.log, device, "python -m meshes_amd.tune_tiles " + " ".join(argv or sys.argv[1:]))


if __name__ == "__main__":
    main()
