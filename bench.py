#!/usr/bin/env python3
"""Benchmark of the per-timestep path on one MI355X.

A "step" is one full Bluebottle timestep: the Lamb iteration loop of u*,
BCs, solvability, Jacobi-PCG pressure solve to the reference's 1e-8 stopping
test, projection, pressure update and Physalis Lamb coefficients, then the
particle motion, all fields resident in HBM.

Default workload: triply periodic sedimentation box 256^3 with 500 spheres
of radius 6 dx (seeded non-overlapping positions, density ratio 2, gravity
in -z, Lamb order 4).  It runs twice, once per form:
  plain          FP64 PCG from phi = 0 (the reference's solver);
  mixed+last     FP32 PCG inside FP64 defect correction, initial guess from
                 the previous solution alone (no network);
  mixed+history  the same, initial guess from the span of the last six
                 solutions (no network);
  mixed+learned  the same, initial guess from the span of the last six
                 solutions and the U-Net's guess, the network trained
                 online on the solver's own solves (its training solves
                 happen in untimed warm-up steps).
The learned form runs first; every form then takes its number of warm-up
steps, so all time the same physical timesteps.
`value` is the mixed+learned form's timesteps/s; the other forms' figures
sit beside it.  Kernel averages cover full launches only: the launches a
batch enqueues past the iteration at which a solve stopped return at once
and their timing samples are dropped.  Prints one JSON line with the throughput, a live roofline
of the dominant kernel of each form (HIP events on the library stream:
`roofline` for the timed form, `roofline_plain` for the plain one) and a
CPU baseline of the FP64 host oracle on a bounded sample.

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
          "spmv_search_kernel_f32": 5, "update_kernel_f32": 6,
          "update_kernel_skip": 8, "update_kernel_pair": 9,
          "update_kernel_f32_skip": 10, "update_kernel_f32_pair": 11}

# Algorithmic bytes per interior cell of each timed kernel (one element per
# array the kernel semantically reads or writes; FP64 = 8 B, FP32 = 4 B),
# plus one byte for the particle mask that the update kernels read when the
# box holds particles (the search kernels read none: the residual is zero
# in particle cells).
KERNEL_BYTES = {
    # update: reads phi, p, r, Ap; writes phi, r
    "update_kernel": 6 * 8,
    # paired phi updates (the default): every other update reads r, Ap and
    # writes r; the next also reads phi and both directions and writes phi
    "update_kernel_skip": 3 * 8,
    "update_kernel_pair": 7 * 8,
    "update_kernel_f32_skip": 3 * 4,
    "update_kernel_f32_pair": 7 * 4,
    # search + SpMV: reads r, p_old; writes p_new, Ap
    "spmv_search_kernel": 4 * 8,
    # the mixed form's inner sweep, FP32 vectors (phi is the correction e)
    "update_kernel_f32": 6 * 4,
    "spmv_search_kernel_f32": 4 * 4,
}
MASK_BYTES = 1

# HBM bytes per interior cell per launch measured by rocprofv3 --pmc
# (FETCH_SIZE x 2 + WRITE_SIZE, the factor 2 from a copy kernel of known
# size) on the 256^3 sedimentation workload with 500 spheres; the commands
# and counter values are in profiles/sediment256_500sph_hbm_traffic.md.
# They apply to that workload (the mask byte included).
TRAFFIC_PER_CELL = {
    "spmv_search_kernel": 34.50,
    "update_kernel": 49.16,
    "update_kernel_skip": 25.15,
    "update_kernel_pair": 57.16,
    "spmv_search_kernel_f32": 16.53,
    "update_kernel_f32_skip": 13.20,
    "update_kernel_f32_pair": 29.34,
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="sediment",
                    choices=["sediment", "channel"],
                    help="sediment: periodic box with spheres (default); "
                         "channel: single-phase channel at Re_tau 180")
    ap.add_argument("--form", default="both",
                    choices=["both", "plain", "mixed-learned", "mixed",
                             "mixed-last", "mixed-history"],
                    help="which forms to time (mixed: FP32-PCG from phi = 0; "
                         "mixed-last: FP32-PCG with the guess from the "
                         "previous solution only; mixed-history: from the "
                         "last six solutions; no network in either)")
    ap.add_argument("--nx", type=int, default=None)
    ap.add_argument("--ny", type=int, default=None)
    ap.add_argument("--nz", type=int, default=None)
    ap.add_argument("--nparts", type=int, default=500)
    ap.add_argument("--radius", type=float, default=6.0)
    ap.add_argument("--seed", type=int, default=20260103)
    ap.add_argument("--train-solves", type=int, default=None,
                    help="solves that feed the network's online training")
    ap.add_argument("--history", type=int, default=None,
                    help="earlier solutions in the span of the history and "
                         "learned forms (default 6)")
    ap.add_argument("--no-cpu", action="store_true",
                    help="skip the CPU oracle baseline")
    ap.add_argument("--spmv-variant", type=int, default=None,
                    help="search kernel: 3 one wavefront per x row where "
                         "the grid allows it, else 2 (default); 2 z-columns "
                         "through LDS planes; 0 direct neighbour loads")
    ap.add_argument("--row-ry", type=int, default=None,
                    help="rows per workgroup of the FP32 row kernel (8 or "
                         "16, default 8)")
    ap.add_argument("--row-zc", type=int, default=None,
                    help="planes per workgroup of the row kernel (multiple "
                         "of 8 up to 64; 0: automatic)")
    ap.add_argument("--row-depth", type=int, default=None,
                    help="planes of loads in flight in the FP32 row kernel "
                         "(2, 3 or 4; default 2)")
    ap.add_argument("--inner-tol", type=float, default=None,
                    help="mixed form: relative fall of the FP32 residual "
                         "norm per inner sweep (pp_inner_tol, default 1e-3)")
    ap.add_argument("--pass-span", type=int, default=None,
                    help="previous timestep's solutions in the Galerkin "
                         "span of the history and learned forms: 1 of the "
                         "same Lamb pass, 2 also of the pass before, 3 also "
                         "of the pass after (library default 3)")
    ap.add_argument("--row-pairs", type=int, default=None,
                    help="1: the FP64 row search loads and stores pairs of "
                         "cells, 1 KB per instruction (default); 0: runs of "
                         "four cells per lane")
    ap.add_argument("--update-variant", type=int, default=None,
                    help="update kernel: 1 one wavefront per x row where "
                         "the grid allows it (default), 0 one tile per "
                         "workgroup")
    ap.add_argument("--urow-zc", type=int, default=None,
                    help="planes per workgroup of the row update (0: "
                         "automatic)")
    ap.add_argument("--urow-depth", type=int, default=None,
                    help="planes of loads in flight in the FP32 row update "
                         "(2 or 4)")
    ap.add_argument("--urow-fp64", type=int, default=None,
                    help="1: the FP64 update uses the row kernel too")
    ap.add_argument("--row-fp64", type=int, default=None,
                    help="1: the FP64 search uses the row kernel too "
                         "(default), 0: the z-column kernel")
    ap.add_argument("--tzb", type=int, default=None,
                    help="tiles per z-column of the search kernel "
                         "(0: automatic)")
    ap.add_argument("--spmv-ntp", type=int, default=None,
                    help="1: FP64 search kernel loads p_old with the "
                         "non-temporal hint (default), 0: ordinary loads")
    ap.add_argument("--spmv-nts", type=int, default=None,
                    help="1: the FP64 z-column search stores p_new and Ap "
                         "with the non-temporal hint (default), 0: ordinary "
                         "stores (the row kernels always use the hint)")
    ap.add_argument("--spmv-nts32", type=int, default=None,
                    help="the same for the FP32 search (default 1)")
    ap.add_argument("--tzb32", type=int, default=None,
                    help="tiles per z-column of the FP32 search "
                         "(0: automatic)")
    ap.add_argument("--defer-phi", type=int, default=None,
                    help="1: the plain form's update kernel adds the "
                         "phi terms of two iterations at once")
    ap.add_argument("--batch", type=int, default=None,
                    help="PCG iterations enqueued per host check")
    ap.add_argument("--predict", type=int, default=None,
                    help="1: the first host check of a sweep after 0.8 of "
                         "its length one timestep earlier (default); 0: "
                         "every batch")
    ap.add_argument("--time-every", type=int, default=8,
                    help="HIP-event timing on one launch in this many of the "
                         "timed kernels (0: no event timing)")
    ap.add_argument("--verbose", action="store_true",
                    help="per-step progress on stderr")
    ap.add_argument("--calib-only", action="store_true",
                    help="run only the stream-copy calibration kernel "
                         "(2^27 doubles per array) and print its GB/s")
    ap.add_argument("--unet", default=None, metavar="N[,N...]",
                    help="run only the learned guess's network: five "
                         "forward passes at each N^3 volume and eight "
                         "training steps at 64^3, timed with device "
                         "synchronisation (for the kernel trace)")
    a = ap.parse_args()
    dflt = (256, 256, 256) if a.workload == "sediment" else (256, 128, 128)
    a.n = tuple(v if v is not None else d
                for v, d in zip((a.nx, a.ny, a.nz), dflt))
    return a


def write_case(a, case, precision, warm):
    ext = {"pp_precision": precision, "pp_warmstart": warm}
    if a.inner_tol is not None:
        ext["pp_inner_tol"] = a.inner_tol
    if a.workload == "channel":
        kw = bbconfig.channel(n=a.n, precision=precision)
        kw["extensions"] = dict(kw.get("extensions") or {}, **ext)
        bbconfig.write_case(case, **kw)
        return
    kw, parts = bbconfig.sediment(n=a.n, nparts=a.nparts, a=a.radius,
                                  seed=a.seed)
    kw["extensions"] = ext
    bbconfig.write_case(case, **kw)
    bbconfig.write_parts(case, parts)


def log(a, *msg):
    if a.verbose:
        print(*msg, file=sys.stderr, flush=True)


def run_form(lib, a, precision, warm, use_net=1, history=4, nwarm=None,
             pass_span=0):
    """Warm-up (and, for the learned guess, the network's training solves),
    then exactly a.steps timed timesteps.  nwarm fixes the warm-up length
    (so that every form times the same physical steps)."""
    warm_mode = warm
    case = tempfile.mkdtemp(prefix="bb_bench_")
    write_case(a, case, precision, warm)
    ctypes.c_int.in_dll(lib, "bb_warm_use_net").value = use_net
    ctypes.c_int.in_dll(lib, "bb_warm_history").value = history
    ctypes.c_int.in_dll(lib, "bb_warm_pass_span").value = pass_span
    tot = ctypes.c_long.in_dll(lib, "bb_pcg_iters_total")
    solves = ctypes.c_long.in_dll(lib, "bb_pcg_solves_total")
    wsolves = ctypes.c_int.in_dll(lib, "bb_warm_solves")
    train = ctypes.c_int.in_dll(lib, "bb_warm_train_solves").value
    sim = bbhip.Simulation(case, a.n)
    tw0 = time.perf_counter()
    sim.setup()
    nw = 0
    while (nw < nwarm) if nwarm is not None else (
            nw < a.warmup or (warm == "LEARNED" and use_net and
                              wsolves.value < train)):
        sim.step()
        nw += 1
        log(a, f"[{precision}/{warm}] warm-up step {nw}: solves "
               f"{wsolves.value}/{train} loss "
               f"{ctypes.c_double.in_dll(lib, 'bb_warm_loss').value:.3e}")
        if nw > a.warmup + 1000:
            raise RuntimeError("network training did not finish")
    sim.sync()
    tw1 = time.perf_counter()
    lib.bb_ktimer_reset()
    ctypes.c_int.in_dll(lib, "bb_time_every").value = max(a.time_every, 1)
    lib.bb_ktimer_enable(1 if a.time_every > 0 else 0)
    it0, sv0 = tot.value, solves.value
    per_step, lamb, resid, solve_iters = [], [], [], []
    slog = (ctypes.c_int * 8192).in_dll(lib, "bb_solve_log")
    sim.sync()
    t0 = time.perf_counter()
    for s in range(a.steps):
        i0, n0 = tot.value, solves.value
        sim.step()
        per_step.append(tot.value - i0)
        if s < 2:
            solve_iters.append([slog[m % 8192]
                                for m in range(n0, solves.value)])
        lamb.append(sim.lamb_stats()[0])
        resid.append(sim.pp_stats()["resid"])
        log(a, f"[{precision}/{warm}] step {s}: PCG iterations "
               f"{per_step[-1]}, Lamb passes {lamb[-1]}")
    sim.sync()
    t1 = time.perf_counter()
    lib.bb_ktimer_enable(0)
    ms = ctypes.c_double()
    cnt = ctypes.c_long()
    ktimes = {}
    dropped = {}
    lib.bb_ktimer_dropped.restype = ctypes.c_long
    for name, kid in KT_IDS.items():
        lib.bb_ktimer_read(kid, ctypes.byref(ms), ctypes.byref(cnt))
        if cnt.value > 0:
            ktimes[name] = (ms.value, cnt.value)
            dropped[name] = lib.bb_ktimer_dropped(kid)
    stats = sim.pp_stats()
    nsolves = solves.value - sv0
    wstat = None
    if use_net and warm_mode == "LEARNED":
        el = ctypes.c_double.in_dll(lib, "bb_warm_err_last_sum").value
        eh = ctypes.c_double.in_dll(lib, "bb_warm_err_hist_sum").value
        eg = ctypes.c_double.in_dll(lib, "bb_warm_err_guess_sum").value
        wstat = {"net_solves": ctypes.c_int.in_dll(
                    lib, "bb_warm_net_solves").value,
                "training_solves": train,
                "energy_error_ratio_last_over_learned":
                    el / eg if eg > 0 else None,
                "energy_error_ratio_history_over_learned":
                    eh / eg if eg > 0 else None}
    sim.teardown()
    shutil.rmtree(case, ignore_errors=True)
    elapsed = t1 - t0
    return {
        "timesteps_per_s": a.steps / elapsed,
        "ms_per_step": 1e3 * elapsed / a.steps,
        "pcg_iters_per_step": sum(per_step) / len(per_step),
        "pcg_iters_per_solve": (tot.value - it0) / max(nsolves, 1),
        "lamb_passes_per_step": sum(lamb) / len(lamb),
        "max_final_resid": max(resid),
        "last_true_resid": stats["true_resid"],
        "warmup_steps": nw,
        "history": history if warm_mode == "LEARNED" else 0,
        "pass_span": pass_span if warm_mode == "LEARNED" else 0,
        "warmup_s": tw1 - tw0,
        "learned_guess": wstat,
        "ktimes": ktimes,
        "samples_dropped_past_convergence": dropped,
        "iters_per_solve_first_timed_steps": solve_iters,
    }


def roofline_of(ktimes, cells, copy_gbs, mask, stream4_gbs=None):
    """Live roofline of the kernel with the most timed time: algorithmic
    bytes per launch over its average launch time (HIP events on the
    library stream)."""
    if not ktimes:
        return None
    dom_name = max(ktimes, key=lambda k: ktimes[k][0])
    kms, kn = ktimes[dom_name]
    avg_s = (kms / 1e3) / kn
    reads_mask = mask and dom_name.startswith("update")
    bpc = KERNEL_BYTES[dom_name] + (MASK_BYTES if reads_mask else 0)
    achieved = bpc * cells / avg_s / 1e9
    tpc = TRAFFIC_PER_CELL.get(dom_name)
    return {
        "kernel": dom_name,
        "bound": "hbm",
        "achieved": achieved,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS,
        "frac_of_measured_copy": achieved / copy_gbs,
        "measured_copy_gbs": copy_gbs,
        "frac_of_stream4": (achieved / stream4_gbs
                            if stream4_gbs and dom_name.endswith("f32")
                            else None),
        "stream4_gbs": stream4_gbs,
        "traffic": tpc * cells if (tpc and mask) else None,
        "bytes_per_cell": bpc,
        "algorithmic_bytes_per_launch": bpc * cells,
        "avg_launch_us": avg_s * 1e6,
        "launches_timed": kn,
    }


def cpu_baseline(a, iters_per_step, lamb_per_step):
    """Time the FP64 host oracle on a bounded sample of the same workload:
    on one oracle instance, truncated timesteps that run every component
    of a timestep (u*, boundary values, solvability, the PCG stopped after
    k iterations, projection, pressure update, Physalis quadrature and
    coefficients, the end-of-step motion, cages and preconditioner): one
    untimed, then (1 Lamb pass, k1), (1 pass, k2) and (2 passes, k1).  They
    give the time of one PCG iteration, of one Lamb pass besides its
    iterations, and of the end-of-step work; a whole timestep is priced
    with the PCG iterations and Lamb passes per timestep measured on the
    GPU's plain FP64 form, whose arithmetic is the oracle's bit for bit
    (the same iteration counts on the parity cases)."""
    import orc  # the checker, used here only as the CPU baseline
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    case = tempfile.mkdtemp(prefix="bb_cpu_")
    write_case(a, case, "PLAIN", "OFF")
    k1, k2 = 2, 22
    o = orc.Oracle(case)
    o.setup()
    o.set_scalar("pcg_continue", 1)
    t = {}
    for key, (passes, k) in (("warm", (1, k1)), ("a", (1, k1)),
                             ("b", (1, k2)), ("c", (2, k1))):
        o.set_scalar("lamb_max_iter", passes)
        o.set_scalar("pp_max_iter", k)
        t0 = time.perf_counter()
        o.lib.orc_step()
        t[key] = time.perf_counter() - t0
    o.teardown()
    shutil.rmtree(case, ignore_errors=True)
    t_iter = max(t["b"] - t["a"], 0.0) / (k2 - k1)
    t_pass = max(t["c"] - t["a"] - k1 * t_iter, 0.0)
    t_end = max(t["a"] - k1 * t_iter - t_pass, 0.0)
    t_step = iters_per_step * t_iter + lamb_per_step * t_pass + t_end
    return {"value": 1.0 / t_step, "unit": "timesteps/s", "cores": cores,
            "kind": "port",
            "sample": f"FP64 C oracle (OpenMP, {cores} threads) on the same "
                      f"workload, truncated timesteps running every "
                      f"component (PCG stopped at {k1} or {k2} iterations, "
                      f"then projection, pressure update, Physalis, motion, "
                      f"cages): 1 pass/{k1} it. {t['a']:.2f} s, 1 pass/{k2} "
                      f"it. {t['b']:.2f} s, 2 passes/{k1} it. {t['c']:.2f} s"
                      f", i.e. {t_iter * 1e3:.1f} ms per PCG iteration, "
                      f"{t_pass:.2f} s per Lamb pass besides its iterations "
                      f"and {t_end:.2f} s of end-of-step work; a timestep "
                      f"priced at the plain GPU form's {iters_per_step:.0f} "
                      f"PCG iterations and {lamb_per_step:.1f} Lamb passes "
                      f"(the oracle's own counts: same arithmetic)"}


def unet_only(lib, vols):
    """The network alone: five forward passes per N^3 volume (random
    inputs in torch device tensors), then eight forward + backward + Adam
    steps at 64^3; prints the mean ms per forward pass and per training
    step."""
    import torch
    lib.bb_unet_create.argtypes = [ctypes.c_ulonglong]
    lib.bb_unet_volume.argtypes = [ctypes.c_int] * 3
    lib.bb_unet_forward.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    lib.bb_unet_backward.argtypes = [ctypes.c_void_p]
    lib.bb_unet_backward.restype = ctypes.c_double
    lib.bb_unet_adam.argtypes = [ctypes.c_float]
    lib.bb_unet_create(7)
    gen = torch.Generator(device="cuda").manual_seed(7)
    out = {}
    for n, reps, train in [(v, 5, False) for v in vols] + [(64, 8, True)]:
        nv = n ** 3
        x = torch.randn(2 * nv, device="cuda", generator=gen)
        t = torch.randn(nv, device="cuda", generator=gen)
        y = torch.zeros(nv, device="cuda")
        lib.bb_unet_volume(n, n, n)
        torch.cuda.synchronize()
        lib.cuda_block()
        t0 = time.perf_counter()
        for _ in range(reps):
            lib.bb_unet_forward(x.data_ptr(), y.data_ptr())
            if train:
                lib.bb_unet_backward(t.data_ptr())
                lib.bb_unet_adam(1e-3)
        lib.cuda_block()
        key = f"train_{n}" if train else f"forward_{n}"
        out[key + "_ms"] = 1e3 * (time.perf_counter() - t0) / reps
    lib.bb_unet_destroy()
    print(json.dumps(out))


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
    lib.bb_stream4_gbs.restype = ctypes.c_double
    lib.bb_stream4_gbs.argtypes = [ctypes.c_long, ctypes.c_int]
    cells = a.n[0] * a.n[1] * a.n[2]
    stream4_gbs = lib.bb_stream4_gbs(cells, 10)
    if a.unet:
        unet_only(lib, [int(v) for v in a.unet.split(",")])
        return
    if a.calib_only:
        print(json.dumps({"stream_copy_gbs": copy_gbs,
                          "elements": 1 << 27,
                          "stream4_gbs": stream4_gbs,
                          "stream4_elements": cells}))
        return
    # the library's default span of earlier timesteps' solutions
    ps_default = ctypes.c_int.in_dll(lib, "bb_warm_pass_span").value
    for name, val in (("bb_spmv_variant", a.spmv_variant),
                      ("bb_spmv_tzb", a.tzb),
                      ("bb_spmv_ntp", a.spmv_ntp),
                      ("bb_defer_phi", a.defer_phi),
                      ("bb_spmv_nts", a.spmv_nts),
                      ("bb_spmv_nts32", a.spmv_nts32),
                      ("bb_spmv_tzb32", a.tzb32),
                      ("bb_row_ry", a.row_ry),
                      ("bb_row_zc", a.row_zc),
                      ("bb_row_depth", a.row_depth),
                      ("bb_row_pairs", a.row_pairs),
                      ("bb_row_fp64", a.row_fp64),
                      ("bb_update_variant", a.update_variant),
                      ("bb_urow_zc", a.urow_zc),
                      ("bb_urow_depth", a.urow_depth),
                      ("bb_urow_fp64", a.urow_fp64),
                      ("bb_pcg_batch", a.batch),
                      ("bb_pcg_predict", a.predict),
                      ("bb_warm_train_solves", a.train_solves)):
        if val is not None:
            ctypes.c_int.in_dll(lib, name).value = val
    # (key, precision, warm start, network, earlier solutions in the span)
    forms = {"both": [("plain", "PLAIN", "OFF", 1, 0),
                      ("mixed_last", "MIXED", "LEARNED", 0, 1),
                      ("mixed_history", "MIXED", "LEARNED", 0, 6),
                      ("mixed_learned", "MIXED", "LEARNED", 1, 6)],
             "plain": [("plain", "PLAIN", "OFF", 1, 0)],
             "mixed-learned": [("mixed_learned", "MIXED", "LEARNED", 1, 6)],
             "mixed": [("mixed", "MIXED", "OFF", 1, 0)],
             "mixed-last": [("mixed_last", "MIXED", "LEARNED", 0, 1)],
             "mixed-history": [("mixed_history", "MIXED", "LEARNED", 0, 6)]
             }[a.form]
    res = {}
    # the learned form first: its warm-up (the network's training solves)
    # sets the warm-up length of the others, so all time the same steps
    order = sorted(forms, key=lambda f: 0 if f[3] and f[2] == "LEARNED"
                   else 1)
    nwarm = None
    for key, prec, warm, use_net, hist in order:
        if a.history is not None and hist > 1:
            hist = a.history
        # the last-solution start (history 1) is that alone; the other
        # learned-mode forms add the previous timestep's solutions
        ps = 0 if hist <= 1 else (a.pass_span if a.pass_span is not None
                                  else ps_default)
        res[key] = run_form(lib, a, prec, warm, use_net, hist, nwarm, ps)
        if nwarm is None:
            nwarm = res[key]["warmup_steps"]
    mask = a.workload == "sediment" and a.nparts > 0
    res = {f[0]: res[f[0]] for f in forms}
    head = res[forms[-1][0]]
    plain = res.get("plain")
    if a.workload == "sediment":
        wl = (f"sedimentation {a.n[0]}x{a.n[1]}x{a.n[2]} triply periodic, "
              f"{a.nparts} spheres a={a.radius:g} dx, rho 2, Lamb order 4, "
              f"PCG to 1e-8; timed form: {forms[-1][0]}")
    else:
        wl = (f"periodic channel {a.n[0]}x{a.n[1]}x{a.n[2]}, Re_tau 180, "
              f"PCG to 1e-8; timed form: {forms[-1][0]}")
    forms_out = {}
    for key, r in res.items():
        forms_out[key] = {k: v for k, v in r.items() if k != "ktimes"}
        forms_out[key]["kernels_avg_us"] = {
            k: 1e3 * v[0] / v[1] for k, v in r["ktimes"].items()}
    out = {
        "metric": METRIC,
        "value": head["timesteps_per_s"],
        "unit": "timesteps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if forms[-1][1] == "PLAIN" else "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": wl,
            "grid": list(a.n),
            "nparts": a.nparts if a.workload == "sediment" else 0,
            "forms": forms_out,
            "speedup_vs_plain": (head["timesteps_per_s"] /
                                 plain["timesteps_per_s"]
                                 if plain and head is not plain else None),
            "spmv_variant": ctypes.c_int.in_dll(lib, "bb_spmv_variant").value,
        },
        "roofline": roofline_of(head["ktimes"], cells, copy_gbs, mask,
                                stream4_gbs),
    }
    if plain is not None and plain is not head:
        out["roofline_plain"] = roofline_of(plain["ktimes"], cells, copy_gbs,
                                            mask, stream4_gbs)
    if not a.no_cpu:
        ref = plain or head
        out["cpu_baseline"] = cpu_baseline(a, ref["pcg_iters_per_step"],
                                           ref["lamb_passes_per_step"])
    print(json.dumps(out))


if __name__ == "__main__":
    main()
