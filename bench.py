#!/usr/bin/env python3
"""Benchmark of the per-timestep flow path on one MI355X.

A "step" is one full Bluebottle timestep (u*, BCs, solvability, Jacobi-PCG
pressure solve to the reference's 1e-8 stopping test, projection, pressure
update, dt) on synthetic input of the periodic-channel workload
(256 x 128 x 128, Re_tau 180, seeded perturbation), all fields resident in
HBM.  Prints one JSON line (rank 0) with the throughput, a live roofline of
the dominant kernel (HIP events on the library stream) and a CPU baseline of
the FP64 host oracle on a bounded sample.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import ctypes
import json
import os
import shutil
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "bluebottle-3.0_amd")
sys.path.insert(0, PKG)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import bbconfig  # noqa: E402
import bbhip  # noqa: E402

METRIC = ("timesteps/sec @256^3+500 spheres (plain FP64 vs mixed+learned); "
          "HBM GB/s")
HBM_PEAK_GBS = 8000.0
# Kernel-timer ids (the KT_* enum of the library).
KT_IDS = {"spmv_search_kernel": 0, "update_kernel": 1,
          "spmv_search_kernel_f32": 5, "update_kernel_f32": 6}

# Algorithmic bytes per interior cell of each timed kernel (one element per
# array the kernel semantically reads or writes; FP64 = 8 B, FP32 = 4 B).
KERNEL_BYTES = {
    # update: reads phi, p, r, Ap; writes phi, r
    "update_kernel": 6 * 8,
    # search + SpMV: reads r, p_old; writes p_new, Ap
    "spmv_search_kernel": 4 * 8,
    # the mixed form's inner sweep, FP32 vectors (phi is the correction e)
    "update_kernel_f32": 6 * 4,
    "spmv_search_kernel_f32": 4 * 4,
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nx", type=int, default=256)
    ap.add_argument("--ny", type=int, default=128)
    ap.add_argument("--nz", type=int, default=128)
    ap.add_argument("--precision", default="PLAIN", choices=["PLAIN", "MIXED"])
    ap.add_argument("--no-cpu", action="store_true",
                    help="skip the CPU oracle baseline")
    ap.add_argument("--spmv-variant", type=int, default=None,
                    help="0: SpMV with direct neighbour loads, 1: LDS-staged "
                         "planes (default)")
    ap.add_argument("--batch", type=int, default=None,
                    help="PCG iterations enqueued per host check")
    ap.add_argument("--time-every", type=int, default=8,
                    help="HIP-event timing on one launch in this many of the "
                         "timed kernels (0: no event timing)")
    ap.add_argument("--calib-only", action="store_true",
                    help="run only the stream-copy calibration kernel "
                         "(2^27 doubles per array) and print its GB/s")
    return ap.parse_args()


def cpu_baseline(case_dir):
    """Time the FP64 host oracle on one timestep of the same workload."""
    import orc  # the checker, used here only as the CPU baseline
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    o = orc.Oracle(case_dir)
    o.setup()
    t0 = time.perf_counter()
    o.step(1)
    t1 = time.perf_counter()
    niter = o.scalar("niter")
    o.teardown()
    return {"value": 1.0 / (t1 - t0), "unit": "timesteps/s", "cores": cores,
            "kind": "port",
            "sample": f"first timestep of the same workload ({int(niter)} PCG "
                      f"iterations), FP64 C oracle with OpenMP on {cores} "
                      f"threads"}


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU that holds the whole "
              f"domain; --gpus {a.gpus} is not supported", file=sys.stderr)
        sys.exit(2)
    lib = bbhip.load()
    lib.bb_stream_copy_gbs.restype = ctypes.c_double
    lib.bb_stream_copy_gbs.argtypes = [ctypes.c_long, ctypes.c_int]
    copy_gbs = lib.bb_stream_copy_gbs(1 << 27, 10)
    if a.calib_only:
        print(json.dumps({"stream_copy_gbs": copy_gbs,
                          "elements": 1 << 27}))
        return
    for name, val in (("bb_spmv_variant", a.spmv_variant),
                      ("bb_pcg_batch", a.batch)):
        if val is not None:
            ctypes.c_int.in_dll(lib, name).value = val
    n = (a.nx, a.ny, a.nz)
    case = tempfile.mkdtemp(prefix="bb_bench_")
    bbconfig.write_case(case, **bbconfig.channel(n=n, precision=a.precision))
    sim = bbhip.Simulation(case, n)
    sim.setup()
    iters = []
    for _ in range(a.warmup):
        sim.step()
        iters.append(sim.pp_stats()["niter"])
    sim.sync()
    lib.bb_ktimer_reset()
    ctypes.c_int.in_dll(lib, "bb_time_every").value = max(a.time_every, 1)
    lib.bb_ktimer_enable(1 if a.time_every > 0 else 0)
    sim.sync()
    t0 = time.perf_counter()
    timed_iters = []
    for _ in range(a.steps):
        sim.step()
        timed_iters.append(sim.pp_stats()["niter"])
    sim.sync()
    t1 = time.perf_counter()
    lib.bb_ktimer_enable(0)
    ms = ctypes.c_double()
    cnt = ctypes.c_long()
    ktimes = {}
    for name, kid in KT_IDS.items():
        lib.bb_ktimer_read(kid, ctypes.byref(ms), ctypes.byref(cnt))
        if cnt.value > 0:
            ktimes[name] = (ms.value, cnt.value)
    stats = sim.pp_stats()
    sim.teardown()

    elapsed = t1 - t0
    cells = a.nx * a.ny * a.nz
    roofline = None
    if ktimes:
        dom_name = max(ktimes, key=lambda k: ktimes[k][0])
        kms, kn = ktimes[dom_name]
        avg_s = (kms / 1e3) / kn
        achieved = KERNEL_BYTES[dom_name] * cells / avg_s / 1e9
        roofline = {
            "kernel": dom_name,
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "frac_of_measured_copy": achieved / copy_gbs,
            "measured_copy_gbs": copy_gbs,
            "traffic": None,
            "bytes_per_cell": KERNEL_BYTES[dom_name],
            "avg_launch_us": avg_s * 1e6,
            "launches_timed": kn,
        }
    out = {
        "metric": METRIC,
        "value": a.steps / elapsed,
        "unit": "timesteps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * elapsed / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if a.precision == "PLAIN" else "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": f"periodic channel {a.nx}x{a.ny}x{a.nz}, Re_tau 180, "
                        f"{a.precision.lower()} Jacobi-PCG to 1e-8",
            "grid": [a.nx, a.ny, a.nz],
            "pp_precision": a.precision,
            "pcg_iters_per_step": sum(timed_iters) / max(len(timed_iters), 1),
            "spmv_variant": ctypes.c_int.in_dll(lib, "bb_spmv_variant").value,
            "last_resid": stats["resid"],
            "last_true_resid": stats["true_resid"],
            "last_corrections": stats["nouter"],
        },
        "roofline": roofline,
        "kernels_avg_us": {k: 1e3 * v[0] / v[1] for k, v in ktimes.items()},
    }
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(case)
    shutil.rmtree(case, ignore_errors=True)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
