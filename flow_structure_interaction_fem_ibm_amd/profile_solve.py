"""Profiling driver: the 256x192x192 stretched-grid solve of bench.py, run a
few times under a profiler, plus a plain device copy of known size used to
calibrate the HBM counters.

  python -m flow_structure_interaction_fem_ibm_amd.profile_solve --solves 2
  python -m flow_structure_interaction_fem_ibm_amd.profile_solve --solves 1 --copy-mib 512
  python -m flow_structure_interaction_fem_ibm_amd.profile_solve --solves 0 --momentum 1

Run it directly after `rocprofv3 ... --` (no launcher in between).
"""
from __future__ import annotations

import argparse
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--solves", type=int, default=2)
    ap.add_argument("--precision", choices=("mixed", "fp64"), default="mixed")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="also copy a buffer of this many MiB device to device 3 times (counter calibration)")
    ap.add_argument("--momentum", type=int, default=0,
                    help="also run the implicit momentum solve of bench.py this many times (after one warm-up)")
    args = ap.parse_args()
    import torch

    import bench
    from flow_structure_interaction_fem_ibm_amd import pp

    dev = torch.device("cuda:0")
    if args.copy_mib > 0:
        n = args.copy_mib * (1 << 20) // 4
        a = torch.rand(n, device=dev)
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(f"copy {args.copy_mib} MiB x3 done", flush=True)
    if args.solves <= 0 and args.momentum <= 0:
        return
    xn, yn, zn, ib, gs, rhs, ex = bench.workload(extras=True)
    s = pp.PressureSolver(xn, yn, zn)
    s.set_geometry(ib, gs)
    if args.momentum > 0:
        s.set_vel_ghosts(ex["coeffD"], 2.0, 1.0)
        r = bench.time_solve_ad(s, dev, reps=args.momentum)
        print(f"momentum: {r['iterations']} iterations {r['ms']:.2f} ms", flush=True)
    d_rhs = torch.from_numpy(rhs).to(dev)
    d_p = torch.zeros(s.shape, dtype=torch.float64, device=dev)
    prec = pp.PP_PREC_MIXED if args.precision == "mixed" else pp.PP_PREC_FP64
    for q in range(args.solves):
        d_p.zero_()
        st = s.solve_device(d_p, d_rhs, precision=prec)
        print(f"solve {q}: {st['status']} {st['outer_iterations']} iterations {st['ms']:.2f} ms", flush=True)


if __name__ == "__main__":
    main()
