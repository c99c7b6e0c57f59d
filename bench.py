#!/usr/bin/env python3
"""Throughput of the mdSpherocyl per-step force path on one MI355X.

Workload (value): N=32768 spherocylinders, lengths Gaussian(3, 0.3)
truncated at +-3 sd, packing fraction 0.20, mu^2=4, Ewald alpha=3/3.5 with
real-space cut-off r_c=3.5 and k_max=10 (reference K input 11), T=1 Nose-Hoover
thermostat, dt=0.002, mixed precision.  One step is the whole device step of
v25/main.cpp:559-622 (thermostat, kicks, drift, forces, observables, list
check and rebuild when due).

Beside it, on the same initial state: the plain FP64 GPU path (steps/s and the
mixed/FP64 ratio), the force/torque/energy error of both against the FP64
oracle, the batched workload (64 replicas, N_i in [512, 8192], packing
fraction in [0.05, 0.30], mu^2 in [1, 6], k_max=7, mixed) as replica-steps/s
and particle-steps/s, and the oracle timed on the host (cpu_baseline).

Prints one JSON line.  Inputs are resident on the device before the timed
region starts; the timed region is msc_run over K steps, which returns after
its stream has drained.
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

# Algorithmic flops per unit, frozen from the operation count of the force
# formulas (never tuned against measurements):
#   reciprocal structure factor: per (particle, +-k pair) one complex product
#     forming e^{-ik.r} from the separable x/y/z factors (6) and three
#     complex-by-real accumulations of mu u_a e^{-ik.r} (12)          -> 18
#   reciprocal back-projection: per (particle, +-k pair) the phase product
#     (6), its product with the folded amplitude A(k) (6), mu u.k (2 with the
#     x/y part hoisted per column), force and field accumulations (7)  -> 21
#   WCA pair (d < 2^{1/6} sigma_ij): segment distance and force/torque    120
#   dipolar real-space pair within r_c: table lerp, force, torques        70
FLOPS = {"recip_struct": 18, "recip_force": 21, "wca": 120, "real_pairs": 70}
# What bounds each kernel of the mixed step: the reciprocal kernels run their
# products on the matrix cores (v_mfma_f32_16x16x32_f16 on operands split into
# two fp16 pieces, three products per FP32-accurate product), priced as FP32
# products against MI355X's dense FP32 matrix peak (157.3 TFLOP/s, equal to
# its FP32 vector peak).  The pair kernels are gathers, priced by their
# algorithmic HBM bytes against 8 TB/s: each particle's records read once
# (rod and dipole, 64 B each), its force and torque read and written (96 B),
# and two 4-byte row entries per interacting pair (WCA inside 2^(1/6) sigma,
# dipole pairs inside r_c).
BOUND = {"recip_struct": "mfma", "recip_force": "mfma", "wca": "hbm", "real_pairs": "hbm"}
PAIR_BYTES_PER_PARTICLE = {"wca": 64 + 96, "real_pairs": 64 + 64 + 96}  # real_pairs: the fused WCA + real-space launch
PAIR_BYTES_PER_PAIR = 8
PEAK_TFLOPS = {"fp32": 157.3, "fp64": 78.6}
HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of a kernel from two rocprofv3 --pmc passes
# (FETCH_SIZE, WRITE_SIZE) scaled by a copy kernel of known size, as
# recorded in the profile summaries
TRAFFIC_FILES = {"recip_force": "profiles/recip_force_mfma_n32768_counters.md",
                 "real_pairs": "profiles/pairs_mixed_n32768_counters.md"}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def n_pm_k(kmax):
    return ((2 * kmax + 1) ** 3 - 1) // 2


def force_split(sysm, rep, sr_cut):
    """Per-routine forces and energies at the current state (lists from
    init_run), as the reference's step calls them."""
    sysm.zero_forces(rep)
    Usc, _, _ = sysm.force_sc(rep, sr_cut)
    UR, _ = sysm.force_dipR(rep)
    UK, _ = sysm.force_dipK(rep)
    F, T = sysm.get_forces(rep)
    return F, T, Usc + UR + UK


def timed_run(sysm, steps, warmup, timing, step0=0):
    if warmup:
        sysm.run(warmup, step0=step0, observe=False)
    sysm.reset_timing()
    sysm.set_timing(timing)
    t0 = time.perf_counter()
    obs = sysm.run(steps, step0=step0 + warmup)
    dt = time.perf_counter() - t0
    sysm.set_timing(False)
    return dt, obs


def autotune(sysm, name, steps=300):
    best, tried = sysm.autotune(steps)
    log("bench: autotune %s: %s" % (name, ", ".join(
        "graph %d fused %d waves %d skin %.2f chunk %d overlap %d %.4f ms" % (
            c["graph"], c["fused_pairs"], c["pair_waves"], c["list_skin"], c["chunk"], c["overlap"], t)
        for c, t in tried)))
    log("bench: autotune %s chose %s" % (name, best))
    return {"choice": best, "steps": steps,
            "tried": [dict(c, ms_per_step=round(t, 5)) for c, t in tried]}


def traffic_from_profile(kernel):
    """HBM bytes per launch of `kernel` recorded by a counter summary (the
    line 'traffic_bytes_per_launch: B'), or None."""
    rel = TRAFFIC_FILES.get(kernel)
    if not rel or not os.path.exists(os.path.join(ROOT, rel)):
        return None, None
    for line in open(os.path.join(ROOT, rel)):
        if line.startswith("traffic_bytes_per_launch:"):
            return int(float(line.split(":", 1)[1])), rel
    return None, None


def device_copy(nbytes):
    """One device-to-device copy of nbytes (the runtime's copy kernel reads
    and writes nbytes each), through the HIP runtime the library uses."""
    path = "/opt/rocm/lib/libamdhip64.so"
    hip = ctypes.CDLL(path if os.path.exists(path) else "libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for p in (a, b):
        if hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(nbytes)):
            raise RuntimeError("hipMalloc failed")
    try:
        if hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) or \
                hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) or hip.hipDeviceSynchronize():
            raise RuntimeError("device copy failed")
    finally:
        hip.hipFree(a)
        hip.hipFree(b)


def host_info():
    """The host the CPU baseline ran on: CPU model (/proc/cpuinfo), the CPUs
    this process may use, and the machine's CPU count."""
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return {"cpu": model, "usable_cpus": usable, "cpu_count": os.cpu_count(),
            "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS")}


def cpu_step_rate(L, reps, budget_s):
    """Steps per second of the oracle's whole reference step (o_run_steps:
    moveTher, moveP, moveX, force_sc + force_dipR + force_dipK, the list check
    and its rebuilds; v25/main.cpp:559-701) over every replica of a workload,
    each replica a run of its own, sampled for about budget_s after one
    untimed step (the first step alone when it takes half the budget)."""
    from oracle import oracle as O
    runs = []
    for d, st in reps:
        S = O.osys_from_replica(d, st)
        P = O.run_params_from_replica(d)
        R = O.CRun()
        if L.o_run_init(S.c(), ctypes.byref(P), ctypes.byref(R), 0):
            raise RuntimeError("oracle run init failed")
        runs.append((S, P, R))

    rebuilds = [0]

    def steps(k0, n):
        obs = np.zeros((n, O.OBS_N))
        t0 = time.perf_counter()
        for S, P, R in runs:
            if L.o_run_steps(S.c(), ctypes.byref(P), ctypes.byref(R), k0, n, O.dptr(obs)):
                raise RuntimeError("oracle step failed")
            rebuilds[0] += int(obs[:, O.OBS_NAMES.index("rebuild")].sum())
        return time.perf_counter() - t0

    t1 = steps(0, 1)
    if t1 >= 0.5 * budget_s:
        n, t = 1, t1  # one step already fills the sample
    else:
        rebuilds[0] = 0
        n = max(1, min(200, int(budget_s / max(t1, 1e-6))))
        t = steps(1, n)
    for _, _, R in runs:
        L.o_run_free(ctypes.byref(R))
    return {"steps_per_s": round(n / t, 5), "steps": n, "seconds": round(t, 2), "list_rebuilds": rebuilds[0],
            "replicas": len(reps), "N": sum(d.N for d, _ in reps)}


def mt19937_state(seed):
    """MT19937 words after init_genrand(seed), position 624 (gsl_rng_set)."""
    mt = np.zeros(624, dtype=np.uint32)
    mt[0] = seed
    for i in range(1, 624):
        prev = int(mt[i - 1])
        mt[i] = (1812433253 * (prev ^ (prev >> 30)) + i) & 0xffffffff
    return mt


def kernel_table(sysm, names):
    out = {}
    for n in names:
        ms, cnt = sysm.kernel_time(n)
        if cnt:
            out[n] = (ms, cnt)
    return out


def cpu_correl_rate(desc, st, n=4096, bar_w=0.01):
    """Pairs per second of the oracle's acc_correl (serial) over the first n
    particles of the workload's state, bars as on the GPU."""
    from oracle import oracle as O
    from mdspherocyl_amd import refio
    S = O.OSys(n, 1, L=tuple(desc.L))
    S.r[:] = st["r"][:n]
    S.ori[:] = st["ori"][:n]
    nb = refio.correl_bars(desc.L[:3], bar_w)
    h = np.zeros((1, 4, nb))
    t0 = time.perf_counter()
    O.lib().o_acc_correl(S.c(), bar_w, nb, O.dptr(h))
    return n * (n - 1) / 2 / (time.perf_counter() - t0)


def cpu_baseline(desc, st, off, nbr, budget_s):
    """The oracle (serial C, FP64) on one force evaluation of the same
    workload: WCA over the same neighbour list, real space over all pairs
    within r_c, reciprocal sum over the same k-vectors."""
    from oracle import oracle as O
    L = O.lib()
    N = desc.N
    S = O.OSys(N, 1, L=tuple(desc.L))
    S.r[:] = st["r"]
    S.ori[:] = st["ori"]
    S.len[:] = st["length"]
    S.dia[:] = st["dia"]
    S.mass[:] = st["mass"]
    S.dpc[:] = st["dpc"]
    S.doc[:] = st["doc"]
    S.dmu[:] = st["dmu"]
    S.set_inertia()
    S.set_dip()
    c = S.c()
    U, W, PT = ctypes.c_double(), ctypes.c_double(), np.zeros(9)
    tab = np.zeros(32768 * 3)
    L.o_tabul(O.dptr(tab), desc.alpha)
    kv = np.zeros((desc.K + 1) ** 3)
    Lb = np.array(desc.L[:3], dtype=np.float64)
    L.o_wave_v(O.dptr(kv), O.dptr(Lb), desc.alpha, desc.K)
    S.F[:] = 0
    S.T[:] = 0
    t0 = time.perf_counter()
    rc = L.o_force_sc(c, desc.sr_cut, O.iptr(off), O.iptr(nbr), ctypes.byref(U), ctypes.byref(W), O.dptr(PT))
    Usc = U.value
    L.o_force_dipR(c, O.dptr(tab), desc.rc, ctypes.byref(U), O.dptr(PT))
    UR = U.value
    L.o_force_dipK(c, desc.K, desc.alpha, O.dptr(kv), ctypes.byref(U), O.dptr(PT))
    UK = U.value
    t = time.perf_counter() - t0
    if rc != 0:
        raise RuntimeError("oracle force_sc reported an overlap on the initial state")
    return t, S.F.copy(), S.T.copy(), Usc + UR + UK


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--fp64-steps", type=int, default=0, help="timed FP64 steps (default: steps / 2)")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="first copy this many MiB device to device (a kernel of known traffic, to scale "
                         "hardware byte counters)")
    ap.add_argument("--cpu-budget", type=float, default=30.0,
                    help="seconds of timed oracle steps over the four workloads")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (replicas would shard across GPUs, "
              "a single system by spatial domains plus one S(k) all-reduce); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    import mdspherocyl_amd as m
    from mdspherocyl_amd import accuracy, configs

    m.load()
    if args.copy_mib > 0:
        device_copy(args.copy_mib << 20)
        log("bench: copied %d MiB device to device" % args.copy_mib)
    t0 = time.perf_counter()
    desc, st = configs.build("c3")[0]
    log("bench: workload N=%d built in %.1f s (box %.3f, alpha %.4f, r_c %.2f, K %d)"
        % (desc.N, time.perf_counter() - t0, desc.L[0], desc.alpha, desc.rc, desc.K))
    kmax = desc.K - 1
    nk = n_pm_k(kmax)

    # --- mixed context (the headline) --------------------------------------
    mix = m.MdSystem([desc], precision="mixed")
    configs.load_into(mix, [(desc, st)])
    mix.init_run()
    Fm, Tm, Um = force_split(mix, 0, desc.sr_cut)
    # the autotuner picks the launch choices (graph, fused pair launch, waves
    # per task) by timing each from this state, then restores it
    tune_c3 = autotune(mix, "N=32768 mixed")

    # --- FP64 context on the same state --------------------------------------
    f64 = m.MdSystem([desc], precision="fp64")
    configs.load_into(f64, [(desc, st)])
    f64.init_run()
    Fd, Td, Ud = force_split(f64, 0, desc.sr_cut)
    off, nbr = f64.nblist(0)
    counts = f64.work_counts(0)
    # a dead evaluation (all-zero or non-finite forces) stops the bench: it
    # would otherwise report perfect agreement
    try:
        err_mix_vs_fp64 = accuracy.force_errors(Fm, Tm, Um, Fd, Td, Ud)
        accuracy.check_reference("mixed F", Fm)
        accuracy.check_reference("mixed T", Tm)
    except accuracy.DegenerateReference as e:
        log("bench: FAILED, degenerate force evaluation: %s" % e)
        sys.exit(3)
    fp64_norms = {"F": float(np.linalg.norm(Fd)), "T": float(np.linalg.norm(Td)), "U": Ud}
    log("bench: mixed vs FP64 GPU at the initial state: F rel L2 %.2e max/rms %.2e max/own %.2e; "
        "T rel L2 %.2e max/rms %.2e max/own %.2e; U %.2e"
        % (err_mix_vs_fp64["F"]["rel_l2"], err_mix_vs_fp64["F"]["max_rel_rms"], err_mix_vs_fp64["F"]["max_rel_own"],
           err_mix_vs_fp64["T"]["rel_l2"], err_mix_vs_fp64["T"]["max_rel_rms"], err_mix_vs_fp64["T"]["max_rel_own"],
           err_mix_vs_fp64["U"]))

    names = ["recip_struct", "recip_force", "real_pairs", "wca", "nlist", "integrate"]
    log("bench: mixed, %d warmup + %d timed steps" % (args.warmup, args.steps))
    # the headline pass runs without event brackets (each bracket leaves the
    # stream idle for several microseconds); the per-kernel times come from a
    # second pass over the next `steps` steps of the same run with brackets
    t_mix, obs = timed_run(mix, args.steps, args.warmup, False)
    counts_mix = mix.work_counts(0)
    rebuilds = int(np.sum(obs[:, 0, m.OBS_NAMES.index("rebuild")]))
    sps_mix = args.steps / t_mix
    log("bench: mixed %.2f steps/s (%.3f ms/step, %d list rebuilds)" % (sps_mix, 1e3 * t_mix / args.steps, rebuilds))
    t_kern, _ = timed_run(mix, args.steps, 0, True, step0=args.warmup + args.steps)
    ktab = kernel_table(mix, names)
    # list rebuilds at their natural rate: a long window of the same run
    n_am = max(1000, args.steps)
    t_am, obs_am = timed_run(mix, n_am, 0, False, step0=args.warmup + 2 * args.steps)
    reb_am = int(np.sum(obs_am[:, 0, m.OBS_NAMES.index("rebuild")]))
    log("bench: mixed over %d steps: %.2f steps/s (%d list rebuilds)" % (n_am, n_am / t_am, reb_am))
    log("bench: per-kernel pass (event brackets on): %.3f ms/step" % (1e3 * t_kern / args.steps))
    for n, (ms, cnt) in sorted(ktab.items(), key=lambda kv: -kv[1][0]):
        log("  %-13s %8.3f ms/launch  (%d launches, %.1f%% of the headline step time)"
            % (n, ms / cnt, cnt, 100 * ms / (1e3 * t_mix)))

    # cluster labels (clust_conf) on the final state, min_s = 0.2: device
    # components plus the host's numbering by the generator's draws
    mix.reset_timing()
    mix.set_timing(True)
    mt, mti = mt19937_state(5), 624
    t0 = time.perf_counter()
    ncl = 3
    for _ in range(ncl):
        labels, nclust, mt, mti, _ = mix.cluster_labels(0, 0.2, mt, mti)
    t_cl = (time.perf_counter() - t0) / ncl
    mix.set_timing(False)
    cl_dev = mix.kernel_time("clusters")
    clusters = {"ms_per_call": round(1e3 * t_cl, 3), "device_ms": round(cl_dev[0] / max(cl_dev[1], 1), 4),
                "clusters": nclust, "largest": int(np.bincount(labels).max()), "min_s": 0.2}
    log("bench: cluster labels %.3f ms per call (device %.3f ms), %d clusters, largest %d"
        % (1e3 * t_cl, clusters["device_ms"], nclust, clusters["largest"]))

    # pair correlations (acc_correl) of the final state, bars of 0.01 as the
    # reference's runs use, n_bars = floor(sqrt(3) Lmin / (2 bar_w))
    from mdspherocyl_amd import refio
    nb = refio.correl_bars(desc.L[:3], 0.01)
    mix.pair_correl(0, 0.01, nb)
    mix.reset_timing()
    mix.set_timing(True)
    t0 = time.perf_counter()
    ncr = 3
    for _ in range(ncr):
        hist, dropped = mix.pair_correl(0, 0.01, nb)
    t_cr = (time.perf_counter() - t0) / ncr
    mix.set_timing(False)
    cr_dev = mix.kernel_time("pair_correl")
    npairs = desc.N * (desc.N - 1) // 2
    correl = {"ms_per_call": round(1e3 * t_cr, 3), "device_ms": round(cr_dev[0] / max(cr_dev[1], 1), 4),
              "n_bars": nb, "bar_w": 0.01, "pairs": npairs, "dropped": dropped,
              "device_pairs_per_s": round(npairs / (cr_dev[0] / max(cr_dev[1], 1) * 1e-3), 1)}
    log("bench: pair correlations %.3f ms per call (device %.3f ms, %.3e pairs/s), %d bars"
        % (1e3 * t_cr, correl["device_ms"], correl["device_pairs_per_s"], nb))

    # orientational order and total moment (oo_params)
    mix.order_params(0)
    mix.reset_timing()
    mix.set_timing(True)
    t0 = time.perf_counter()
    for _ in range(10):
        G1, G2, Mtot = mix.order_params(0)
    t_oo = (time.perf_counter() - t0) / 10
    mix.set_timing(False)
    oo_dev = mix.kernel_time("order_params")
    order = {"ms_per_call": round(1e3 * t_oo, 4), "device_ms": round(oo_dev[0] / max(oo_dev[1], 1), 4),
             "G1": G1, "G2": G2, "M": [float(m) for m in Mtot]}
    log("bench: order parameters %.4f ms per call (device %.4f ms), G1 %.4f G2 %.4f"
        % (1e3 * t_oo, order["device_ms"], G1, G2))

    k64 = args.fp64_steps or args.steps
    log("bench: FP64, %d warmup + %d timed steps" % (min(args.warmup, 10), k64))
    t_d, _ = timed_run(f64, k64, min(args.warmup, 10), False)
    sps_d = k64 / t_d
    log("bench: FP64 %.2f steps/s; mixed/FP64 = %.2f" % (sps_d, sps_mix / sps_d))
    t_dam, obs_dam = timed_run(f64, n_am, 0, False, step0=min(args.warmup, 10) + k64)
    reb_dam = int(np.sum(obs_dam[:, 0, m.OBS_NAMES.index("rebuild")]))
    log("bench: FP64 over %d steps: %.2f steps/s (%d list rebuilds); mixed/FP64 = %.2f"
        % (n_am, n_am / t_dam, reb_dam, t_dam / t_am))
    amortized = {"steps": n_am, "mixed_steps_per_s": round(n_am / t_am, 3), "mixed_list_rebuilds": reb_am,
                 "fp64_steps_per_s": round(n_am / t_dam, 3), "fp64_list_rebuilds": reb_dam,
                 "mixed_over_fp64": round(t_dam / t_am, 3)}
    # FP64 with the WCA over the contact rows, as the mixed path walks them
    f64.set_contact_rows(True)
    t_dc, _ = timed_run(f64, k64, min(args.warmup, 10), False)
    f64.set_contact_rows(False)
    sps_dc = k64 / t_dc
    log("bench: FP64 on contact rows %.2f steps/s; mixed/that = %.2f" % (sps_dc, sps_mix / sps_dc))
    mix.close()
    f64.close()

    # --- roofline of every force kernel; `roofline` is the dominant one ----------
    # (the kernel with the most time per step; the fused launch "real_pairs"
    # holds both the WCA and the real-space pairs)
    wca_pairs, real_pairs = counts_mix[0], counts_mix[1]
    fused = "wca" not in ktab or ktab["wca"][1] == 0
    pair_flop = {"wca": FLOPS["wca"] * wca_pairs,
                 "real_pairs": FLOPS["real_pairs"] * real_pairs + (FLOPS["wca"] * wca_pairs if fused else 0)}
    pair_bytes = {"wca": PAIR_BYTES_PER_PARTICLE["wca"] * desc.N + PAIR_BYTES_PER_PAIR * wca_pairs,
                  "real_pairs": (PAIR_BYTES_PER_PARTICLE["real_pairs"] * desc.N
                                 + PAIR_BYTES_PER_PAIR * (real_pairs + (wca_pairs if fused else 0)))}

    def roofline_of(name):
        ms, cnt = ktab[name]
        sec = ms / cnt * 1e-3
        traffic, traffic_src = traffic_from_profile(name)
        if BOUND[name] == "mfma":
            flop = FLOPS[name] * desc.N * nk
            ach = flop / sec / 1e12
            r = {"kernel": name, "bound": "mfma", "achieved": round(ach, 3), "peak": PEAK_TFLOPS["fp32"],
                 "unit": "TFLOP/s", "frac": round(ach / PEAK_TFLOPS["fp32"], 4), "flop_per_launch": flop}
        else:
            nbytes = pair_bytes[name]
            ach = nbytes / sec / 1e9
            r = {"kernel": name, "bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                 "frac": round(ach / HBM_PEAK_GBS, 4), "bytes_per_launch": nbytes,
                 "flop_per_launch": pair_flop[name],
                 "fp64_tflops": round(pair_flop[name] / sec / 1e12, 3)}
        r.update({"traffic": traffic, "traffic_source": traffic_src, "ms_per_launch": round(ms / cnt, 4)})
        return r

    force_kernels = [n for n in ktab if n in FLOPS and ktab[n][1] > 0]
    roofline_by_kernel = {n: roofline_of(n) for n in force_kernels}
    dom = max(force_kernels, key=lambda n: ktab[n][0])
    roofline = roofline_by_kernel[dom]
    for n, r in roofline_by_kernel.items():
        log("bench: roofline %-12s %8.4f ms/launch  %s %.3g %s of %.4g (frac %.4f)%s"
            % (n, r["ms_per_launch"], r["bound"], r["achieved"], r["unit"], r["peak"], r["frac"],
               "  <- dominant" if n == dom else ""))

    # --- batched replicas ---------------------------------------------------------
    batched = None
    reps_c4 = None
    if not args.no_batched:
        t0 = time.perf_counter()
        reps = reps_c4 = configs.build("c4")
        ntot = sum(d.N for d, _ in reps)
        log("bench: batched workload, %d replicas, sum N = %d, built in %.1f s"
            % (len(reps), ntot, time.perf_counter() - t0))
        bat = m.MdSystem([d for d, _ in reps], precision="mixed")
        configs.load_into(bat, reps)
        bat.init_run()
        tune_c4 = autotune(bat, "64 replicas mixed", steps=50)
        kb = max(args.steps // 2, 10)
        t_b, _ = timed_run(bat, kb, min(args.warmup, 10), False)
        repeats = bat.chunk_repeats()
        bat.close()
        sps_b = kb / t_b
        pps_b = sps_b * ntot
        pps_single = sps_mix * desc.N
        batched = {"replicas": len(reps), "sum_N": ntot, "steps": kb, "steps_per_s": round(sps_b, 3),
                   "replica_steps_per_s": round(sps_b * len(reps), 2), "particle_steps_per_s": round(pps_b, 1),
                   "vs_single_particle_steps": round(pps_b / pps_single, 3), "chunk_repeats": repeats,
                   "tuning": tune_c4}
        log("bench: batched %.2f steps/s = %.1f replica-steps/s; particle-steps/s %.3e = %.2f x single system"
            % (sps_b, sps_b * len(reps), pps_b, pps_b / pps_single))

    # --- oracle on the host ----------------------------------------------------------
    cpu = None
    err_oracle = None
    if not args.no_cpu:
        from oracle import oracle as O
        Lomp = O.lib_omp()
        threads = int(Lomp.o_threads())
        host = host_info()
        log("bench: host %s, %d usable CPUs, oracle timing build on %d threads"
            % (host["cpu"], host["usable_cpus"], threads))
        per = {}
        for name in ("c1", "c2", "c3", "c4"):
            reps = [(desc, st)] if name == "c3" else (reps_c4 if name == "c4" and reps_c4 else configs.build(name))
            per[name] = cpu_step_rate(Lomp, reps, args.cpu_budget / 4)
            log("bench: oracle full step %s (%d replicas, sum N %d): %.4g steps/s over %d steps on %d threads"
                % (name, per[name]["replicas"], per[name]["N"], per[name]["steps_per_s"], per[name]["steps"], threads))
        log("bench: oracle force evaluation on the host (serial FP64)")
        t_cpu, Fo, To, Uo = cpu_baseline(desc, st, off, nbr, args.cpu_budget)
        cpu = {"value": per["c3"]["steps_per_s"], "unit": "steps/s", "cores": threads, "kind": "port",
               "sample": "%d whole reference step(s) (integrator, WCA over the reference list, real space "
                         "within r_c=3.5, reciprocal k_max=10, list check; %d list rebuilds) of the N=32768 "
                         "workload, the oracle's OpenMP timing build on %d threads, %.1f s"
                         % (per["c3"]["steps"], per["c3"]["list_rebuilds"], threads, per["c3"]["seconds"]),
               "host": host, "workloads": per,
               "serial_force_evals_per_s": round(1.0 / t_cpu, 5),
               "serial_sample": "one force evaluation of the N=32768 workload by the serial oracle, %.1f s" % t_cpu}
        err_oracle = {"mixed": accuracy.force_errors(Fm, Tm, Um, Fo, To, Uo),
                      "fp64": accuracy.force_errors(Fd, Td, Ud, Fo, To, Uo)}
        for k in ("mixed", "fp64"):
            e = err_oracle[k]
            log("bench: %s vs oracle: F rel L2 %.2e max/rms %.2e max/own %.2e; T rel L2 %.2e max/rms %.2e "
                "max/own %.2e; U %.2e" % (k, e["F"]["rel_l2"], e["F"]["max_rel_rms"], e["F"]["max_rel_own"],
                                          e["T"]["rel_l2"], e["T"]["max_rel_rms"], e["T"]["max_rel_own"], e["U"]))
        correl["oracle_pairs_per_s"] = round(cpu_correl_rate(desc, st), 1)
        log("bench: oracle acc_correl %.3e pairs/s (serial, first 4096 particles)" % correl["oracle_pairs_per_s"])

    line = {
        "metric": "steps/s @N=32768 mixed vs FP64 GPU (force err≤1e-5); batched replica-steps/s",
        "value": round(sps_mix, 3),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "list_rebuilds": rebuilds,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * t_mix / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "N=32768 spherocylinders, L~N(3,0.3), phi=0.20, mu^2=4, alpha=3/3.5, r_c=3.5, "
                               "k_max=10, mixed precision", "N": desc.N, "box": round(desc.L[0], 4),
                   "k_pairs": nk, "wca_pairs": counts[0], "real_pairs": counts[1], "list_pairs": counts[3]},
        "fp64_steps_per_s": round(sps_d, 3),
        "mixed_over_fp64": round(sps_mix / sps_d, 3),
        "fp64_contact_rows_steps_per_s": round(sps_dc, 3),
        "mixed_over_fp64_contact_rows": round(sps_mix / sps_dc, 3),
        "amortized": amortized,
        "tuning": tune_c3,
        "err_mixed_vs_fp64": err_mix_vs_fp64,
        "fp64_force_norms": fp64_norms,
        "err_vs_oracle": err_oracle,
        "batched": batched,
        "cluster_labels": clusters,
        "pair_correl": correl,
        "order_params": order,
        "kernels_ms_per_launch": {n: round(v[0] / v[1], 4) for n, v in ktab.items()},
        "roofline": roofline,
        "roofline_by_kernel": roofline_by_kernel,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
