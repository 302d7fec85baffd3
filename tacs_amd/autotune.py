"""Tuner of the kernels that come in several launch variants with identical
results: the sliced ELL block product (slots per unrolled pass, non-temporal
value loads), the multi right-hand-side block product (blocks per lane per
trip, non-temporal loads) and the solid Jacobian (rows kernel: wavefronts per
workgroup and elements per geometry chunk; or the element-block form).  For
every workload below it runs the library's own timers (Mat.autotune,
Mat.autotuneSpMM, Assembler.autotune: HIP events on the library stream) and
writes the fastest variant per (kernel, shape, size class) to
tacs_amd/tuning_table.json, which `import tacs_amd` enters into the library.
The reference's counterpart is the block-size-specialised kernel choice of
BCSRMat::initBlockImpl (src/bpmat/BCSRMat.cpp:1028).

    python -m tacs_amd.autotune [--out tacs_amd/tuning_table.json] [--quick]
"""
import argparse
import json
import os
import sys
import time

SIZE_CLASS_ROWS = "0: < 2^15 block rows, 1: < 2^19, 2: >= 2^19"


def _device_name():
    try:
        import torch
        return torch.cuda.get_device_name(0)
    except Exception:  # noqa: BLE001
        return "unknown"


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                                  "tuning_table.json"))
    ap.add_argument("--quick", action="store_true", help="small meshes only")
    args = ap.parse_args(argv)
    from tacs_amd import TACS
    from tacs_amd.problems import ShellProblem, SolidProblem

    TACS.resetTuning()  # time every variant against the built-in defaults
    L = TACS.lib()
    best = {}  # (kernel, shape, size_class) -> (ms, variant, workload, times)

    def enter(kernel, shape, rows, times, workload):
        sc = L.tacs_tune_size_class(rows)
        v = min(range(len(times)), key=lambda i: times[i])
        key = (kernel, shape, sc)
        # several workloads of one size class: keep the one timed on the
        # largest (longest-running) mesh
        if key not in best or rows > best[key][0]:
            best[key] = (rows, v, workload, times)
        print(f"  kernel {kernel} shape {shape} class {sc}: variant {v} "
              f"({', '.join(f'{t:.4f}' for t in times)} ms)", flush=True)

    solids = [(2, 20), (2, 64), (3, 10), (3, 24)]
    shells = [128, 512]
    if not args.quick:
        solids += [(2, 128), (3, 40)]
        shells += [1024]
    for order, n in solids:
        t0 = time.perf_counter()
        p = SolidProblem(n, order=order)
        p.assemble()
        b, nr, nnz = p.mat.getInfo()
        name = f"hex{8 if order == 2 else 27} {n}^3"
        print(f"{name}: {nr} block rows, {nnz} blocks (setup {time.perf_counter() - t0:.1f} s)",
              flush=True)
        r = p.mat.autotune(reps=20)
        for kind in range(3):
            enter(TACS.TUNE_SELL_SPMV, 30 + kind, nr, r["ms"][kind], name)
        jt = p.assembler.autotune(p.mat, reps=3)
        if jt:
            enter(TACS.TUNE_JACOBIAN, order, nr, jt, name)
        if order == 3:
            for nrhs in (4, 8, 16):
                for prec in (32, 64):
                    t = p.mat.autotuneSpMM(nrhs, prec=prec, reps=10)
                    enter(TACS.TUNE_SPMM, (100 if prec == 32 else 0) + nrhs, nr, t,
                          f"{name}, {nrhs} right-hand sides")
        del p
    for n in shells:
        t0 = time.perf_counter()
        p = ShellProblem(n)
        p.assemble()
        b, nr, nnz = p.mat.getInfo()
        name = f"Quad4 {n}^2"
        print(f"{name}: {nr} block rows (setup {time.perf_counter() - t0:.1f} s)", flush=True)
        r = p.mat.autotune(reps=20)
        for kind in range(3):
            enter(TACS.TUNE_SELL_SPMV, 60 + kind, nr, r["ms"][kind], name)
        del p

    names = {TACS.TUNE_SELL_SPMV: "sell_spmv", TACS.TUNE_SPMM: "bsr3_spmm",
             TACS.TUNE_JACOBIAN: "solid_jacobian"}
    entries = [{"kernel": names[k], "kernel_id": k, "shape": sh, "size_class": sc,
                "variant": v, "workload": wl, "block_rows": rows,
                "ms": [round(t, 5) for t in times]}
               for (k, sh, sc), (rows, v, wl, times) in sorted(best.items())]
    table = {"device": _device_name(), "size_classes": SIZE_CLASS_ROWS,
             "command": "python -m tacs_amd.autotune" + (" --quick" if args.quick else ""),
             "entries": entries}
    with open(args.out, "w") as f:
        json.dump(table, f, indent=1)
    print(f"wrote {len(entries)} entries to {args.out}")
    TACS.loadTuningTable(args.out)
    return 0


if __name__ == "__main__":
    sys.exit(main())
