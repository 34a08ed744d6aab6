#!/usr/bin/env python3
"""Benchmark of the PIM-Ewald force + dipole step on one MI355X.

A step is one engforce evaluation (BMHFTD pair forces, polarizable-ion SCF
with Ewald charge+dipole sums, final Ewald forces/stress and EFG at every
site) of the 24,576-atom polarizable silica-like melt (8192 Si + 16384 O,
rc 12 A, kmax 24, SCF tolerance 1e-7), on positions already resident in HBM.
Successive steps take successive frames of a seeded random walk (0.0125 A
per component per frame, the thermal displacement of O at 3000 K in 1 fs),
so the dipole extrapolation and SCF behave as in an MD run.  Each path runs
W warmup steps, K steps with every launch bracketed by HIP events (the
per-kernel table, untimed), then K timed steps with events around the
launches of the dominant kernel only (its live average launch time gives the
roofline figure).

Prints ONE JSON line.  The plain FP64 path and the mixed-precision path are
both timed; ``value`` is the mixed path's atom-steps/s (FP64 when the mixed
path is not requested), with its force error against FP64 stated.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "atom-steps/s PIM-Ewald (mixed vs FP64, 1 GPU) at stated force err; grad evals/s"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", type=int, default=5, choices=[2, 3, 5])
    ap.add_argument("--precision", default="both", choices=["fp64", "mixed", "both"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the oracle CPU baseline")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the batched force-matching gradient leg")
    ap.add_argument("--batch-evals", type=int, default=5,
                    help="timed loss+gradient evaluations of the batch leg")
    ap.add_argument("--tuning", default=None, metavar="SLICES,SPLIT",
                    help="fix the mixed path's matrix-core launch shape (apply K slices, "
                         "structure-factor split atoms) instead of autotuning it")
    ap.add_argument("--mixed-scf", default="pcg", choices=["jacobi", "pcg"],
                    help="SCF solver of the mixed path (the FP64 path always runs the "
                         "reference's Jacobi iteration); with pcg the mixed path is also "
                         "timed with the Jacobi iteration, reported beside the headline")
    ap.add_argument("--event-every", type=int, default=4,
                    help="HIP events on the dominant kernel's launches of every N-th timed "
                         "step (each event pair stalls the queue a few microseconds)")
    ap.add_argument("--quiet", action="store_true")
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="run one plain device copy of this size first (calibrates HBM "
                         "byte counters of a profiler against a known transfer)")
    return ap.parse_args()


def log(args, *a):
    if not args.quiet:
        print(*a, file=sys.stderr, flush=True)


def frames_for(sys_, nframes, seed=20240601, amp=0.0125):
    import numpy as np
    rng = np.random.default_rng(seed)
    out = np.empty((nframes, 3, sys_.natm))
    r = sys_.r.copy()
    for f in range(nframes):
        out[f] = r.T
        r = r + amp * rng.standard_normal(r.shape)
    return out


def profile_traffic(kernel):
    """HBM bytes per launch of a kernel event scope, as measured with
    rocprofv3 FETCH_SIZE / WRITE_SIZE passes (copy-kernel calibrated) and
    recorded in a committed summary under profiles/ as the line
    'traffic_bytes_per_launch <scope> <bytes>'; (None, None) if absent."""
    import glob
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*traffic*.md"))):
        for line in open(path):
            f = line.split()
            if len(f) == 3 and f[0] == "traffic_bytes_per_launch" and f[1] == kernel:
                return float(f[2]), os.path.relpath(path, ROOT)
    return None, None


def time_path(args, sys_, precision, frames_dev, scf=None, tuning=None, table=True):
    """Times K steps of one path (precision, SCF solver); with table, K more
    untimed steps first fill the per-kernel table and the dominant kernel's
    roofline is measured on the timed steps."""
    import copy
    import torch
    from mdff_amd import params as P
    from mdff_amd.field import ForceField
    from mdff_amd import flops as FL
    p = copy.deepcopy(sys_.params)
    p.precision = precision
    if scf is None:
        scf = args.mixed_scf if precision == P.PREC_MIXED else "jacobi"
    if scf == "pcg":
        p.scf_solver = P.SCF_PCG
    n = sys_.natm
    force = torch.zeros((3, n), dtype=torch.float64, device="cuda")
    if tuning:
        pass
    elif precision == P.PREC_MIXED and args.tuning:
        sl, sp = (int(x) for x in args.tuning.split(","))
        tuning = {"apply_slices": sl, "split_atoms": sp}
    elif precision == P.PREC_MIXED:
        # launch shape of the matrix-core reciprocal sums, picked on a
        # separate context (its dipole history is not the timed one's)
        ft = ForceField(p, sys_.itype, sys_.A)
        tuning = ft.autotune(frames_dev[0], force, itime=1, reps=2)
        ft.close()
    ff = ForceField(p, sys_.itype, sys_.A)
    if tuning and not ff.set_mixed_tuning(tuning["apply_slices"], tuning["split_atoms"]):
        raise SystemExit(f"bench: launch shape {tuning} rejected by the library")
    W, K = args.warmup, args.steps
    for w in range(W):
        ff.engforce_device(frames_dev[w], force, itime=w + 1)
    stats = {}
    name = None
    if table:
        # per-kernel table: every launch bracketed by HIP events, K untimed
        # steps
        ff.profile(True)
        for k in range(K):
            ff.engforce_device(frames_dev[W + k], force, itime=W + k + 1)
        torch.cuda.synchronize()
        stats = ff.profile_stats()
        name = max(stats.items(), key=lambda kv: kv[1][0])[0]
        # timed region: the next K frames of the trajectory, events around
        # the launches of the dominant kernel on every event_every-th step
        # (every step has the same launch mix)
        ff.profile(True, only=name)
    every = max(1, args.event_every)
    nscf, nind, rmsd, scf_tol = [], [], [], []
    sampled = []
    b = W + K
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(K):
        if table:
            ff.profile_select(name if k % every == 0 else None)
        en = ff.engforce_device(frames_dev[b + k], force, itime=b + k + 1)
        nscf.append(en.n_scf)
        nind.append(en.n_es_calls - 2)  # induced-dipole Ewald calls of the step
        rmsd.append(en.rmsd)
        scf_tol.append(en.scf_tol)
        if k % every == 0:
            sampled.append(k)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    if table:
        ms, cnt = ff.profile_stats()[name]
        ff.profile(False)
    # last frame forces for the error statement
    f_last = force.clone()
    dt = (t1 - t0) / K
    nk = ff.num_kvectors()
    ff.close()
    # roofline of the dominant kernel from its live launches: achieved =
    # algorithmic flops per launch (per-(atom, k) weights of mdff_amd/flops.py
    # times N K) / the average HIP-event launch time
    roof = None
    flop = None
    if table:
        avg_s = ms / cnt * 1e-3
        per_unit = FL.KERNEL_FLOPS_PER_UNIT.get(name)
        nsc = sum(nind[k] for k in sampled)
        # the static and SCF calls' field (both paths) and the mixed path's
        # dipole-only structure factors run over the polarizable atoms (no
        # species here has a static dipole)
        npol = int(sum(1 for t in sys_.itype if p.ldip_polar[t - 1]))
        if per_unit:
            # the k_apply64<field> launches are the static and SCF calls
            flop = per_unit * (npol if name == "k_apply64<field>" else n) * nk
            peak, bound, pipe = FL.PEAK_FP64_TFLOPS, "valu", "fp64 vector ALU"
        elif name in ("mx_rho", "mx_apply"):
            flop = FL.kspace_algorithmic_flops(name, n, nk, len(sampled), nsc, p.doefg,
                                               n_dip=npol, n_fld=npol) / cnt
            executed = FL.mixed_kspace_flops(name, n, list(p.kES), len(sampled), nsc, p.doefg,
                                             n_dip=npol, n_fld=npol) / cnt
            peak, bound = FL.PEAK_FP16_MFMA_TFLOPS, "mfma"
            pipe = "fp16 matrix core (operands split in fp16 hi + lo, 3 products)"
    if flop:
        achieved = flop / avg_s / 1e12
        traffic, source = profile_traffic(name)
        roof = {"bound": bound, "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                "frac": achieved / peak, "traffic": traffic, "kernel": name,
                "avg_launch_us": avg_s * 1e6, "launches": cnt, "sampled_steps": len(sampled),
                "flops_per_launch": flop, "pipe": pipe}
        if bound == "mfma":
            roof["executed_mfma_flops_per_launch"] = executed
            roof["executed_mfma_frac"] = executed / avg_s / 1e12 / peak
        if source:
            roof["traffic_source"] = source
    return {"ms_per_step": dt * 1e3, "atom_steps_per_s": n / dt, "n_scf": nscf,
            "rmsd_max": max(rmsd), "scf_tol_max": max(scf_tol), "scf": scf,
            "kernels": {k: {"total_ms": v[0], "launches": v[1]} for k, v in stats.items()},
            "roofline": roof, "force": f_last,
            "tuning": ({"apply_slices": tuning["apply_slices"],
                        "split_atoms": tuning["split_atoms"]} if tuning else None)}


BATCH_SCF = 20  # fixed SCF iterations of the captured loss+gradient graph


def batch_leg(args):
    """Batched force matching: 64 polarizable silica-like configurations
    with atom counts uniform in [96, 1536] (kmax 8, cutoffs min(10 A, L/2)),
    target forces from the model at hidden parameters (2% seeded
    perturbation) plus N(0, 0.05 eV/A) noise.  One loss + adjoint gradient
    evaluation (BMHFTD parameters, charges, polarizabilities) is one replayed
    hipGraph with BATCH_SCF SCF iterations; timed for FP64 and mixed, with
    the mixed loss, gradient and forces compared against FP64."""
    import numpy as np
    import torch
    from mdff_amd import configs as C
    from mdff_amd import params as P
    from mdff_amd.batch import Batch, autotune, get_theta, set_theta
    t0 = time.perf_counter()
    systems = C.batch(nconf=64, precision=P.PREC_FP64)
    natm = sum(s.natm for s in systems)
    log(args, f"[bench] batch of {len(systems)} configurations, {natm} atoms, generated in "
              f"{time.perf_counter() - t0:.1f}s")
    pos = torch.tensor(np.concatenate([s.r for s in systems], axis=0).T.copy(),
                       dtype=torch.float64, device="cuda")
    b64 = Batch.from_systems(systems, precision=P.PREC_FP64)
    b32 = Batch.from_systems(systems, precision=P.PREC_MIXED)
    b64.set_positions(pos)
    b32.set_positions(pos)
    rng = np.random.default_rng(41)
    th0 = get_theta(b64.model)
    hidden = th0 * (1.0 + 0.02 * rng.standard_normal(th0.shape))
    b64.set_model(set_theta(b64.model, hidden))
    f_hidden, _ = b64.forces(BATCH_SCF)
    b64.set_model(set_theta(b64.model, th0))
    noise = torch.tensor(0.05 * rng.standard_normal(tuple(f_hidden.shape)), device="cuda")
    target = (f_hidden + noise).contiguous()
    # launch shapes by the autotuner: FP32 matrix-core K slices and split
    # size, FP64 reciprocal-apply lanes per atom
    tune32 = autotune(b32, target, BATCH_SCF, reps=2, use_cache=False)
    tune64 = autotune(b64, target, BATCH_SCF, reps=2, use_cache=False)
    out = {"nconf": len(systems), "natm": natm, "scf_iters": BATCH_SCF,
           "evals": args.batch_evals, "tuning": {"mixed": tune32["best"],
                                                 "fp64": tune64["best"]}}
    ref = {}
    for key, b in (("fp64", b64), ("mixed", b32)):
        loss, grad, forces = b.loss_grad(target, BATCH_SCF, want_forces=True)
        ref[key] = (loss, grad, forces)
        b.graph(True)
        for _ in range(2):
            b.loss_grad(target, BATCH_SCF)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(args.batch_evals):
            b.loss_grad(target, BATCH_SCF)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t1) / args.batch_evals
        b.graph(False)
        out[key] = {"grad_evals_per_s": 1.0 / dt, "ms_per_grad_eval": dt * 1e3,
                    "atom_evals_per_s": natm / dt}
        log(args, f"[bench] batch {key}: {dt * 1e3:.2f} ms per loss+gradient")
    l64, g64, f64 = ref["fp64"]
    l32, g32, f32 = ref["mixed"]
    out["mixed_vs_fp64"] = {
        "speedup": out["mixed"]["grad_evals_per_s"] / out["fp64"]["grad_evals_per_s"],
        "loss_rel_err": abs(l32 - l64) / abs(l64),
        "grad_max_rel_err": float(np.max(np.abs(g32 - g64)) / np.max(np.abs(g64))),
        "force_rel_rms_err": float(torch.sqrt(((f32 - f64) ** 2).mean())
                                   / torch.sqrt((f64 ** 2).mean())),
    }
    b64.close()
    b32.close()
    return out


def cpu_baseline(args, sys_, n_ind_per_step):
    """Oracle (plain C, 1 thread) on a bounded sample of the same workload.
    Timed in full: the BMHFTD pair sum.  Each of the step's three kinds of
    Ewald call (static: charges, field; induced: dipoles, field; final:
    charges + dipoles, field, EFG, forces, stress) is timed with all atoms
    and the full real-space sum at kmax 1, and on the first quarter of the
    atoms at kmax 3 and kmax 6: the k-space cost is linear in atoms x k
    vectors, so the difference of the last two gives the cost per (atom, k),
    scaled to all atoms and the workload's k vectors.  Per step = BMHFTD +
    static + n_ind induced + final."""
    import copy
    import numpy as np
    from oracle import oracle as O
    from mdff_amd import params as P
    O.build()
    n = sys_.natm
    p = copy.deepcopy(sys_.params)
    kfull = P.num_kvectors(list(p.kES))
    q = np.array([p.qch[t - 1] for t in sys_.itype])
    mu = np.where((sys_.itype == 2)[:, None],
                  0.05 * np.random.default_rng(7).standard_normal((n, 3)), 0.0)
    measured = 0.0
    # short range: BMHFTD only
    psr = copy.deepcopy(p)
    psr.lcoulomb = 0
    t0 = time.perf_counter()
    O.Oracle(psr, sys_.itype, sys_.A).engforce(sys_.r)
    t_sr = time.perf_counter() - t0
    measured += t_sr
    kinds = {"static": (q, np.zeros((n, 3)), [1, 0, 0], dict(do_efg=False, do_forces=False,
                                                             do_stress=False)),
             "induced": (np.zeros(n), mu, [0, 0, 1], dict(do_efg=False, do_forces=False,
                                                          do_stress=False, damp_ind=False)),
             "final": (q, mu, [1, 1, 1], dict(do_efg=bool(p.doefg)))}
    per_call = {}
    m = n // 4
    for kind, (qq, mm, task, kw) in kinds.items():
        t = []
        for km, na in ((3, m), (6, m), (1, n)):
            pk = copy.deepcopy(p)
            pk.kES[:] = [km, km, km]
            pk.natm = na
            t0 = time.perf_counter()
            O.multipole_es(pk, sys_.itype[:na], sys_.A, sys_.r[:na], qq[:na], mm[:na], task, **kw)
            t.append((P.num_kvectors([km] * 3), time.perf_counter() - t0))
            measured += t[-1][1]
        (k1, t1), (k2, t2), (k0, t0) = t
        per_atom_k = max(t2 - t1, 0.0) / ((k2 - k1) * m)
        real = max(t0 - per_atom_k * k0 * n, 0.0)
        per_call[kind] = real + per_atom_k * kfull * n
    step = t_sr + per_call["static"] + n_ind_per_step * per_call["induced"] + per_call["final"]
    # the host's cores: the reference decomposes atoms and k points over its
    # MPI ranks, modelled as `cores` equal shares of the step; their
    # concurrent throughput is measured with that many worker processes
    # running the induced-dipole k-space sample at once against one alone
    cores = max(1, min(CPU_WORKERS, os.cpu_count() or 1))
    pk = copy.deepcopy(p)
    pk.kES[:] = [6, 6, 6]
    pk.natm = m
    eff, t_one, t_all = _parallel_efficiency(pk, sys_.itype[:m], sys_.A, sys_.r[:m], mu[:m], cores)
    measured += t_one + t_all
    one_core = n / step
    return {"value": one_core * cores * eff, "unit": "atom-steps/s", "cores": cores,
            "kind": "port",
            "sample": (f"oracle (C) on the same {n}-atom configuration: on one core the BMHFTD "
                       f"pair sum, and each kind of Ewald call (static, induced, final) with "
                       f"all atoms at kmax 1 (real space) and with {m} atoms at kmax 3 and 6 "
                       f"(k-space cost per atom and k vector), scaled to {n} atoms and "
                       f"{kfull} k vectors; per step BMHFTD + static + {n_ind_per_step:.2f} "
                       f"induced + final call; times {cores} cores at the efficiency of "
                       f"{cores} concurrent worker processes each running the induced call "
                       f"on {m} atoms at kmax 6 ({t_all:.2f} s) against one alone "
                       f"({t_one:.2f} s)"),
            "one_core_value": one_core,
            "parallel_efficiency": eff,
            "seconds_per_step_estimate": step / (cores * eff),
            "seconds_measured": measured}


CPU_WORKERS = 16  # the host's CPU share per GPU

_CPU_WORKER = r"""
import sys, time
import numpy as np
sys.path.insert(0, sys.argv[1])
from oracle import oracle as O
from mdff_amd import params as P
d = np.load(sys.argv[2])
p = P.MdffParams.from_buffer_copy(d["params"].tobytes())
n = len(d["itype"])
t0 = time.perf_counter()
O.multipole_es(p, d["itype"], d["A"], d["r"], np.zeros(n), d["mu"], [0, 0, 1], damp_ind=False,
               do_efg=False, do_forces=False, do_stress=False)
print(time.perf_counter() - t0, flush=True)
"""


def _parallel_efficiency(p, itype, A, r, mu, cores):
    """Time of one oracle induced-dipole call alone and of `cores` copies in
    concurrent worker processes (CPU only, each timing its own call);
    efficiency = alone / slowest concurrent copy."""
    import subprocess
    import tempfile
    import numpy as np
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "sample.npz")
        np.savez(path, params=np.frombuffer(bytes(p), dtype=np.uint8), itype=itype, A=A, r=r,
                 mu=mu)
        cmd = [sys.executable, "-c", _CPU_WORKER, ROOT, path]

        def run(k):
            # the longest of the workers' own timings of the call (process
            # start-up excluded)
            procs = [subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE)
                     for _ in range(k)]
            worst = 0.0
            for pr in procs:
                out, err = pr.communicate()
                if pr.returncode != 0:
                    raise RuntimeError(f"cpu baseline worker failed: {err.decode()[-400:]}")
                worst = max(worst, float(out.decode().split()[-1]))
            return worst
        t_one = run(1)
        t_all = run(cores)
    return min(1.0, t_one / t_all), t_one, t_all


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    from mdff_amd import configs as C
    from mdff_amd import params as P
    if not torch.cuda.is_available():
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    if args.copy_mb:
        nbytes = args.copy_mb * (1 << 20)
        src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(args, f"[bench] calibration copy of {args.copy_mb} MiB done")
        del src, dst
    t0 = time.perf_counter()
    sys_ = C.config(args.workload)
    log(args, f"[bench] workload {sys_.name} n={sys_.natm} generated in "
              f"{time.perf_counter() - t0:.1f}s")
    frames = frames_for(sys_, args.warmup + 2 * args.steps)
    frames_dev = torch.tensor(frames, dtype=torch.float64, device="cuda")
    precisions = {"fp64": [P.PREC_FP64], "mixed": [P.PREC_MIXED],
                  "both": [P.PREC_FP64, P.PREC_MIXED]}[args.precision]
    res = {}
    for prec in precisions:
        key = "fp64" if prec == P.PREC_FP64 else "mixed"
        log(args, f"[bench] timing {key} ...")
        res[key] = time_path(args, sys_, prec, frames_dev)
        log(args, f"[bench] {key}: {res[key]['ms_per_step']:.2f} ms/step, "
                  f"{res[key]['atom_steps_per_s']:.4g} atom-steps/s, n_scf {res[key]['n_scf']}")
    if "mixed" in res and "fp64" in res and args.mixed_scf == "pcg":
        # the speedup at equal algorithm: the mixed path with the
        # reference's Jacobi iteration, and the FP64 path with the
        # conjugate-gradient SCF
        log(args, "[bench] timing mixed with the Jacobi SCF ...")
        res["mixed_jacobi"] = time_path(args, sys_, P.PREC_MIXED, frames_dev, scf="jacobi",
                                        tuning=res["mixed"]["tuning"], table=False)
        log(args, "[bench] timing fp64 with the conjugate-gradient SCF ...")
        res["fp64_pcg"] = time_path(args, sys_, P.PREC_FP64, frames_dev, scf="pcg", table=False)
    head = res.get("mixed", res.get("fp64"))

    def ferr(key):
        a = res[key]["force"]
        b = res["fp64"]["force"]
        return float(torch.sqrt(((a - b) ** 2).mean()) / torch.sqrt((b ** 2).mean()))
    err = ferr("mixed") if "mixed" in res and "fp64" in res else None
    cpu = None
    if not args.no_cpu:
        # the reference's Jacobi iteration: one induced-dipole call per SCF
        # iteration
        n_ind = float(np.mean(res["fp64" if "fp64" in res else "mixed"]["n_scf"]))
        cpu = cpu_baseline(args, sys_, n_ind)
        log(args, f"[bench] cpu baseline {cpu['value']:.4g} atom-steps/s "
                  f"(measured {cpu['seconds_measured']:.1f}s)")
    p = sys_.params
    out = {
        "metric": METRIC,
        "value": head["atom_steps_per_s"],
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        # the mixed path: FP32 pair and real-space arithmetic, reciprocal
        # sums on the fp16 matrix cores with every operand split in fp16
        # hi + lo (22-bit significands, 3 products, FP32 accumulation),
        # FP64 reductions and SCF residual
        "dtype": "f16x2-split mfma + f32 + f64" if "mixed" in res else "f64",
        "data": "synthetic",
        "config": {
            "workload": (f"{sys_.name}: polarizable silica-like melt, {sys_.natm} atoms, "
                         f"L={sys_.info['L']:.4f} A, rc {p.cutlongrange} A, kmax {p.kES[0]}, "
                         f"alpha {p.alphaES:.5f}, SCF tol {p.conv_tol_ind:g}, EFG at all sites; "
                         f"random placement with minimum separations Si-Si "
                         f"{C.PIM_DMIN[(1, 1)]} A, Si-O {C.PIM_DMIN[(1, 2)]} A, O-O "
                         f"{C.PIM_DMIN[(2, 2)]} A (a common 1.2 A floor lets O-O pairs "
                         f"polarize catastrophically)"
                         if p.ldip_polar[1] else sys_.name),
            "natm": sys_.natm,
            "kvectors": P.num_kvectors(list(p.kES)),
            "path": "mixed" if "mixed" in res else "fp64",
            "scf_solver": {"mixed": args.mixed_scf, "fp64": "jacobi"},
        },
        "roofline": head["roofline"],
        "cpu_baseline": cpu,
    }
    for key in ("fp64", "mixed"):
        if key in res:
            r = res[key]
            out[key] = {"atom_steps_per_s": r["atom_steps_per_s"], "ms_per_step": r["ms_per_step"],
                        "mean_n_scf": float(np.mean(r["n_scf"])), "scf_solver": r["scf"],
                        "scf_rmsd_max": r["rmsd_max"], "scf_tol_max": r["scf_tol_max"],
                        "roofline": r["roofline"], "kernels": r["kernels"]}
            if r["tuning"]:
                out[key]["tuning"] = r["tuning"]
    if err is not None:
        out["mixed_vs_fp64"] = {"speedup": res["fp64"]["ms_per_step"] / res["mixed"]["ms_per_step"],
                                "force_rel_rms_err": err}
    for key in ("mixed_jacobi", "fp64_pcg"):
        if key not in res:
            continue
        r = res[key]
        out[key] = {"atom_steps_per_s": r["atom_steps_per_s"], "ms_per_step": r["ms_per_step"],
                    "mean_n_scf": float(np.mean(r["n_scf"])), "scf_rmsd_max": r["rmsd_max"],
                    "scf_tol_max": r["scf_tol_max"], "force_rel_rms_err": ferr(key)}
    if "mixed_jacobi" in res and "fp64_pcg" in res:
        # mixed against FP64 with the same SCF algorithm
        out["mixed_vs_fp64"]["speedup_jacobi"] = (res["fp64"]["ms_per_step"]
                                                  / res["mixed_jacobi"]["ms_per_step"])
        out["mixed_vs_fp64"]["speedup_pcg"] = (res["fp64_pcg"]["ms_per_step"]
                                               / res["mixed"]["ms_per_step"])
    if not args.no_batch:
        out["batch"] = batch_leg(args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
