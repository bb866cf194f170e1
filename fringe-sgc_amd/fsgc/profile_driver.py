"""Driver for rocprofv3 runs of the counting path.

    python fringe-sgc_amd/fsgc/profile_driver.py [--workload c2] [--copy-mib 1024]

First a device-to-device copy of a known size (torch, `copy_`, the calibration
kernel for FETCH_SIZE/WRITE_SIZE), then one untimed count (index build and
automorphism sum) and one measured count of the workload, on the graph renumbered
by degree when the workload asks for it (as bench.py does). Prints what it ran.
"""
import argparse
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="c2")
    ap.add_argument("--copy-mib", type=int, default=1024)
    ap.add_argument("--relabel", default="auto", choices=["auto", "none", "desc", "asc"],
                    help="degree relabelling before the counts (auto: the workload's choice in bench.py)")
    args = ap.parse_args()
    import torch
    import fsgc
    from bench import WORKLOADS
    n = args.copy_mib * (1 << 20) // 4
    x = torch.arange(n, dtype=torch.int32, device="cuda")
    y = torch.empty_like(x)
    y.copy_(x)
    torch.cuda.synchronize()
    print(f"calibration copy: {args.copy_mib} MiB read + {args.copy_mib} MiB written", flush=True)
    del x, y
    W = WORKLOADS[args.workload]
    fsgc.init(0)
    g = fsgc.Graph.generate(W["spec"])
    order = W.get("relabel", "none") if args.relabel == "auto" else args.relabel
    if order != "none":
        g0 = g
        g, _, ms = g0.relabel(order)
        g0.free()
        print(f"relabelled by degree ({order}) in {ms:.1f} ms", flush=True)
    mo, _ = fsgc.preprocess(*W["motif"])
    r0 = g.count(mo)
    print(f"warm count: index {r0.index_ms:.1f} ms kernel {r0.kernel_ms:.1f} ms", flush=True)
    r = g.count(mo)
    print(f"count: kernel {r.kernel_ms:.1f} ms, alg_bytes {r.alg_bytes:.4g}, core matches {r.core_matches}, "
          f"exact {r.exact}", flush=True)


if __name__ == "__main__":
    main()
