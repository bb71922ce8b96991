"""Times one library call shape: a Clements mesh of N modes and L layers, B vectors, one op.

    python -m meshes_amd.profile_case --op fwdgrad --X mzi --N 1024 --L 1024 --B 2048 --steps 5

Runs ``--warmup`` untimed calls, then ``--steps`` calls bracketed by HIP events on the current
stream, and prints one JSON line (ms per call).  Meant as the program under ``rocprofv3`` for
kernel traces and counter passes of a given size; the tile is the library's choice unless
``--tile w,v,waves,chunk`` is given.  ``--op hvp`` times an exact Hessian-vector product of the
target loss: fwddiff along a random direction q, then meshprop_hvp.
"""
from __future__ import annotations

import argparse
import json

import torch

from . import _lib as Lb
from .tune_tiles import OPS, Case


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--op", choices=list(OPS) + ["hvp"], default="fwdgrad")
    ap.add_argument("--X", choices=["mzi", "sym", "orth", "mzi_generic"], default="mzi")
    ap.add_argument("--N", type=int, default=256)
    ap.add_argument("--L", type=int, default=0, help="layers (default N)")
    ap.add_argument("--B", type=int, default=4096)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--tile", type=str, default="")
    ap.add_argument("--generic-kernels", action="store_true",
                    help="run the general forward and adjoint kernels instead of the NN-specialised ones")
    args = ap.parse_args(argv)
    L = args.L or args.N
    tile = tuple(int(t) for t in args.tile.split(",")) if args.tile else None
    hvp = args.op == "hvp"
    op = Lb.OP_FWDDIFF if hvp else OPS[args.op]
    case = Case(op, args.X, args.N, L, args.B)
    if args.generic_kernels:
        case.extra_flags = Lb.GENERIC_KERNELS
    nb = Lb.workspace_size(op, args.X, args.N, L, args.B, False, tile)
    ws = torch.empty(max(nb, 256), dtype=torch.uint8, device="cuda")
    call = lambda: case.run(tile, ws, nb)  # noqa: E731
    if hvp:
        torch.manual_seed(0)
        case.dp.normal_()  # the direction q
        nh = Lb.hvp_workspace_size(args.X, args.N, L, args.B, False)
        wh = torch.empty(max(nh, 256), dtype=torch.uint8, device="cuda")
        hv = torch.zeros(case.p.numel(), dtype=torch.float32, device="cuda")

        def call():
            case.run(tile, ws, nb)
            Lb.hvp(args.X, args.N, L, args.B, False, case.lens, case.shifts, case.p, case.dp, case.ldp, case.s,
                   case.lds, case.out, case.dout, None, None, case.y, None, hv, args.N, wh, nh)
    for _ in range(args.warmup):
        call()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(args.steps):
        call()
    b.record()
    b.synchronize()
    ms = a.elapsed_time(b) / max(1, args.steps)
    out = {"op": args.op, "X": args.X, "N": args.N, "L": L, "B": args.B,
           "tile": list(Lb.resolve_tile(op, args.X, args.N, args.B, tile)) if not (args.X == "mzi_generic" and
                                                                                  op == Lb.OP_FWDGRAD) else tile,
           "generic_kernels": bool(args.generic_kernels), "ms_per_call": ms, "updates_per_s": args.B * case.n_active / (ms * 1e-3),
           "device": torch.cuda.get_device_name(0)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
