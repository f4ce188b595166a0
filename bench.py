"""Benchmark: direct-SCF J/K (Fock) builds per second on one MI355X.

Workload: 20-water cluster, 60 atoms, even-tempered s/p basis plus one d
shell on O (N = 440 Cartesian functions), Schwarz tau = 1e-12; mixed
precision runs quartets with Q_ab Q_cd < 1e-5 in FP32.  One step = one full
J+K build (all quartet classes + deterministic reduction) for a fixed
synthetic closed-shell density P = 2 C C^T (100 orthonormal random
orbitals), inputs resident in HBM.

Prints one JSON line (rank 0).  `value` is FP64 Fock builds/s; the line also
carries the mixed-precision rate, quartet rates, the two-electron energy
difference mixed vs FP64, the roofline of the dominant kernel (HIP events on
the launch stream) and the CPU baseline (the C oracle, a bounded sample of
the same quartets on the host cores).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6   # MI355X FP64 (vector and matrix)


def measured_traffic(kernel_ls):
    """HBM bytes per launch of the dominant kernel from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE summary (profiles/), if one exists."""
    import re
    name = "k_eri_class_" + "".join(map(str, kernel_ls)) + "_h2o20.md"
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    m = re.search(r"HBM traffic per launch = ([0-9.eE+]+) bytes", open(path).read())
    return float(m.group(1)) if m else None


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="target host time of the CPU-baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-scf", action="store_true", help="skip the FP64/mixed SCF energy check")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the 256-molecule batched SCF measurement")
    ap.add_argument("--batch-size", type=int, default=256)
    ap.add_argument("--gradient-only", action="store_true",
                    help="only the (H2O)10 SCF + three FP64 nuclear gradients (for kernel tracing)")
    ap.add_argument("--builds-only", action="store_true",
                    help="autotune, then only W + K J/K builds (for kernel tracing)")
    ap.add_argument("--precision", type=int, default=0,
                    help="precision of the --builds-only builds (0 FP64, 1 mixed)")
    ap.add_argument("--streams", type=int, default=-1,
                    help="side streams of the --builds-only builds (-1: the library's default; "
                         "0: every launch in order, so a kernel trace shows each kernel alone)")
    ap.add_argument("--only-batch", action="store_true",
                    help="run only the batched SCF workload and print its JSON")
    ap.add_argument("--only-published", action="store_true",
                    help="run only the reference's published cc-pVDZ molecules")
    ap.add_argument("--mixed-sweep", action="store_true",
                    help="only sweep the FP32 quartets' primitive threshold (mixed builds)")
    ap.add_argument("--workload", type=int, default=1, help="workload index (1 = 20 waters)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (for HBM counter scaling)")
    return ap.parse_args()


def copy_calibration(mib: int):
    """A plain device-to-device copy of known size, to scale FETCH_SIZE /
    WRITE_SIZE counter readings of the ERI kernels."""
    import torch
    n = mib * (1 << 20) // 8
    a = torch.arange(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": n * 8, "read_bytes": n * 8, "write_bytes": n * 8}))


def synthetic_density(n: int, nocc: int, seed: int = 0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    C, _ = np.linalg.qr(rng.standard_normal((n, nocc)))
    return 2.0 * C @ C.T


def cpu_baseline(basis, mol, work, seconds: float):
    """The reference's per-component path on the host cores: the oracle's
    restatement of rys.eri_rys (core.f90:917-1002: one contracted Cartesian
    component quartet per call, every primitive quartet with 5 Rys nodes and
    the 2-D recurrence + transfer per node and direction), with the nodes
    from fitted tables (orc_rys5_fast, the evaluation-cost class of root5's
    piecewise fits), on a seeded uniform sample of the surviving shell
    quartets' component quartets.  Timed with 8 threads (the reference's
    Pool(processes=8), hf.py:45) and with 1; extrapolated to one J/K build
    by the unique function-quartet count (digestion not counted, which
    favours the CPU)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    shells = mol.shells()
    npgf = mol.cgf_arrays()[0]
    rng = np.random.default_rng(1234)
    classes = [w for w in work if w["shellq"] > 0]
    weights = np.array([w["shellq"] for w in classes], dtype=np.float64)
    weights /= weights.sum()
    pairs = {pc: basis.pair_list(pc)[:2] for pc in range(6)}
    lists = {w["class"]: basis.quartet_list(w["class"])[0] for w in classes}

    def sample_quartets(nshellq):
        quads = []
        for _ in range(nshellq):
            w = classes[rng.choice(len(classes), p=weights)]
            la, lb, lc, ld = w["ls"]
            pcb, pck = la * (la + 1) // 2 + lb, lc * (lc + 1) // 2 + ld
            n_all = lists[w["class"]]
            cum = np.cumsum(n_all)
            q = int(rng.integers(0, cum[-1]))
            i = int(np.searchsorted(cum, q, side="right"))
            j = q - (int(cum[i - 1]) if i else 0)
            a, b = pairs[pcb][0][i], pairs[pcb][1][i]
            c, d = pairs[pck][0][j], pairs[pck][1][j]
            for fa in range(shells[a].nfunc):
                for fb in range(shells[b].nfunc):
                    for fc in range(shells[c].nfunc):
                        for fd in range(shells[d].nfunc):
                            quads.append((shells[a].offset + fa, shells[b].offset + fb,
                                          shells[c].offset + fc, shells[d].offset + fd))
        return np.array(quads, dtype=np.int32)

    threads = min(8, os.cpu_count() or 1)   # Pool(processes=8), hf.py:45
    O.lib().orc_rys5_fast_init()

    def rate(qs, nthreads):
        t0 = time.perf_counter()
        O.eri_list(mol, qs, nthreads, fast_roots=True)
        return len(qs) / max(time.perf_counter() - t0, 1e-9)

    # calibrate, then size each timed sample to its share of `seconds`
    qcal = sample_quartets(200)
    per_shellq = len(qcal) / 200.0
    r8 = rate(qcal, threads)
    qs8 = sample_quartets(max(200, min(int(r8 * seconds * 0.4 / per_shellq), 400000)))
    qs8 = qs8[rng.permutation(len(qs8))]
    t0 = time.perf_counter()
    O.eri_list(mol, qs8, threads, fast_roots=True)
    dt8 = time.perf_counter() - t0
    rate8 = len(qs8) / dt8
    qs1 = qs8[: max(200, int(rate8 / threads * seconds * 0.4))]
    t0 = time.perf_counter()
    O.eri_list(mol, qs1, 1, fast_roots=True)
    dt1 = time.perf_counter() - t0
    rate1 = len(qs1) / dt1
    primq1 = int(np.prod(npgf[qs1].astype(np.int64), axis=1).sum())
    us_primq = dt1 / primq1 * 1e6
    us_root = O.lib().orc_time_rys5_fast(1_000_000, 40.0)   # 1e6 calls -> s = us per call
    funcq = sum(w["funcq"] for w in work)
    return {"value": rate8 / funcq, "unit": "Fock builds/s", "cores": threads, "kind": "port",
            "sample": f"{len(qs8)} contracted component quartets from a seeded uniform sample of "
                      f"surviving shell quartets of the same build: oracle restatement of "
                      f"rys.eri_rys (5 nodes per primitive quartet, nodes from fitted tables of "
                      f"root5's cost class), {threads} threads {dt8:.1f} s, 1 thread on the first "
                      f"{len(qs1)} {dt1:.1f} s; {rate8:.3e} quartets/s ({rate1:.3e} on 1 thread) "
                      f"extrapolated by {funcq} unique function quartets per build (digestion not "
                      f"counted)",
            "quartets_per_s": rate8,
            "single_thread": {"quartets_per_s": rate1, "builds_per_s": rate1 / funcq},
            "us_per_primitive_quartet_1thread": us_primq,
            "us_per_root_evaluation": us_root,
            "root_share": us_root / us_primq}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is design-only "
                                   "(partition of each class's quartet ranges across GPUs, then an all-reduce of "
                                   "X and Y), not run; --gpus must be 1"}))
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.only_batch:
        print(json.dumps({"batch": batch_workload(args.batch_size)}))
        return
    if args.only_published:
        print(json.dumps({"published": published_workload()}))
        return
    if args.mixed_sweep:
        print(json.dumps({"mixed_sweep": mixed_sweep(args.steps)}))
        return
    if args.gradient_only:
        print(json.dumps({"gradient_calls": gradient_calls(3)}))
        return
    import torch
    from segfault_py_amd import engine, metrics, workloads as W

    mol = W.config_molecule(args.workload)
    tau, tau_mixed = 1e-12, 1e-5
    t0 = time.perf_counter()
    basis = engine.Basis(mol).screen(tau, tau_mixed)
    setup_s = time.perf_counter() - t0
    nocc = int(mol.numbers.sum()) // 2
    P = torch.tensor(synthetic_density(mol.nbf, nocc), dtype=torch.float64, device="cuda")
    J = torch.empty_like(P)
    K = torch.empty_like(P)
    if args.builds_only:
        # W warm-up and K timed FP64 builds, nothing else: each build starts
        # with the density-scaling kernel, which splits a kernel trace per build
        prec = args.precision
        if args.streams >= 0:
            basis.set_streams(args.streams)
        basis.autotune(P, prec)
        for _ in range(args.warmup):
            basis.jk(P, prec, out=(J, K))
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(args.steps):
            basis.jk(P, prec, out=(J, K))
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t) / args.steps * 1e3
        print(json.dumps({"builds_only": {"workload": mol.nbf, "precision": prec,
                                          "streams": args.streams, "ms_per_build": ms,
                                          "builds": args.steps}}))
        return
    # flops count the primitive quartets the FP64 kernels evaluate
    work = metrics.class_work(basis, mol, evaluated=basis.prim_quartet_counts(0),
                              far=basis.far_prim_quartet_counts(0))

    basis.autotune(P, 0)
    tuned = basis.autotune(P, 1)[0]      # shapes of every slot after both tunings

    def timed(precision, shapes=tuned):
        basis.set_tuning(shapes)
        for _ in range(args.warmup):
            basis.jk(P, precision, out=(J, K))
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(args.steps):
            basis.jk(P, precision, out=(J, K))
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / args.steps

    t64 = timed(0)
    E64 = float(0.5 * (P * (J - 0.5 * K)).sum())

    # the same FP64 build replayed from a captured hipGraph (sf_jk_graph),
    # as the SCF driver runs it every iteration
    def timed_graph(precision):
        basis.set_tuning(tuned)
        for _ in range(max(1, args.warmup)):
            basis.jk_graph(P, (J, K), precision)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(args.steps):
            basis.jk_graph(P, (J, K), precision)
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / args.steps

    t64g = timed_graph(0)
    E64g = float(0.5 * (P * (J - 0.5 * K)).sum())
    tmix = timed(1)
    Emix = float(0.5 * (P * (J - 0.5 * K)).sum())
    # every small-quartet segment in FP32 arithmetic (no FP64 fall-back) ...
    fp32_only = tuned.copy()
    fp32_only[:, 2, 2] = 0
    tmix32 = timed(1, fp32_only)
    Emix32 = float(0.5 * (P * (J - 0.5 * K)).sum())
    # ... and also without the primitive screening of those quartets
    basis.set_prim_tau(0.0)
    tmix_np = timed(1, fp32_only)
    Emix_np = float(0.5 * (P * (J - 0.5 * K)).sum())
    basis.set_prim_tau(-1.0)

    # per-kernel times from HIP events on the launch stream
    basis.set_tuning(tuned)
    _, _, ms = basis.jk_profile(P, 0, out=(J, K))
    kern = [(ms[2 * w["class"]], w) for w in work if w["shellq"] > 0]
    dom_ms, dom = max(kern, key=lambda x: x[0])
    top = sorted(kern, key=lambda x: -x[0])[:6]
    kernel_ms = {"eri_" + "".join(map(str, w["ls"])): round(t, 4) for t, w in top}
    kernel_ms["reduce+finalize"] = round(float(ms[engine.SLOT_REDUCE]), 4)
    kernel_ms["scale_density"] = round(float(ms[engine.SLOT_SCALE]), 4)
    # ceiling of the mixed build: its in-order launch groups with the FP32
    # segments' groups (the FP32-arithmetic launches, and for split classes
    # their shared digestion launch) taken as free, against the in-order
    # FP64 build
    _, _, ms_mix = basis.jk_profile(P, 1, out=(J, K))
    fp32_groups = float(sum(ms_mix[2 * c + 1] for c in range(engine.NUM_QUARTET_CLASSES)))
    mixed_bound = {"speedup_bound": float(sum(ms)) / max(1e-9, float(sum(ms_mix)) - fp32_groups),
                   "fp64_inorder_ms": round(float(sum(ms)), 4),
                   "mixed_inorder_ms": round(float(sum(ms_mix)), 4),
                   "fp32_segment_groups_ms": round(fp32_groups, 4),
                   "note": "Amdahl bound: in-order FP64 build / (in-order mixed build minus the "
                           "FP32 segments' launch groups), i.e. the mixed speed-up if the "
                           "FP32-arithmetic launches cost nothing; a loose ceiling, since those "
                           "launches also do the FP64-pinned node, prefactor and digestion work"}
    # the McMurchie-Davidson engine (MFMA E.R.E) on the same build, per class
    basis.set_engine(1)
    basis.jk_profile(P, 0, out=(J, K))
    _, _, ms_mmd = basis.jk_profile(P, 0, out=(J, K))
    basis.set_engine(0)
    engines = {"eri_" + "".join(map(str, w["ls"])): {"rys_ms": round(float(ms[2 * w["class"]]), 4),
                                                      "mmd_mfma_ms": round(float(ms_mmd[2 * w["class"]]), 4)}
               for w in work if w["shellq"] > 0}
    engines["build_ms"] = {"rys": round(float(sum(ms)), 3), "mmd_mfma": round(float(sum(ms_mmd)), 3)}
    achieved = dom["flops"] / (dom_ms * 1e-3) / 1e12
    # every class against the same ceiling: its kernels' time in this build
    # (integrals + digestion) and its algorithmic flops
    class_roofline = {"eri_" + "".join(map(str, w["ls"])): {
        "ms": round(float(ms[2 * w["class"]]), 4),
        "shell_quartets": int(w["shellq"]), "prim_quartets": int(w["primq"]),
        "tflops": round(w["flops"] / (ms[2 * w["class"]] * 1e-3) / 1e12, 2),
        "frac": round(w["flops"] / (ms[2 * w["class"]] * 1e-3) / 1e12 / FP64_PEAK_TFLOPS, 3)}
        for w in work if w["shellq"] > 0 and ms[2 * w["class"]] > 0}
    total_flops = sum(w["flops"] for w in work)
    shellq = sum(w["shellq"] for w in work)
    funcq = sum(w["funcq"] for w in work)
    primq = sum(w["primq"] for w in work)

    line = {
        "metric": "J/K Fock builds/s & ERI quartets/s: mixed vs FP64 GPU at |dE|<1e-6 Eh",
        "value": 1.0 / t64,
        "unit": "Fock builds/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t64 * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"(H2O)20 direct-SCF J+K build, 60 atoms, N={mol.nbf}, "
                               "even-tempered s/p + d(O), Schwarz 1e-12",
                   "nbf": mol.nbf, "tau": tau, "tau_mixed": tau_mixed, "parallelism": "single-gpu"},
        "fp64_builds_per_s": 1.0 / t64,
        "fp64_builds_per_s_graph_replay": {"builds_per_s": 1.0 / t64g,
                                           "dE_2e_vs_eager": abs(E64g - E64)},
        "mixed_builds_per_s": 1.0 / tmix,
        "mixed_speedup": t64 / tmix,
        "mixed_speedup_bound": mixed_bound,
        "dE_2e_mixed_vs_fp64": abs(Emix - E64),
        "mixed_fp64_arith_classes": [int(c) for c in np.nonzero(tuned[:, 2, 2])[0]],
        "mixed_fp32_arith_only": {"builds_per_s": 1.0 / tmix32, "speedup": t64 / tmix32,
                                  "dE_2e_vs_fp64": abs(Emix32 - E64)},
        "mixed_fp32_arith_no_prim_screen": {"builds_per_s": 1.0 / tmix_np, "speedup": t64 / tmix_np,
                                            "dE_2e_vs_fp64": abs(Emix_np - E64)},
        "quartets_per_s_fp64": funcq / t64,
        "shell_quartets_per_s_fp64": shellq / t64,
        "prim_quartets_per_s_fp64": primq / t64,
        "quartets_per_build": {"function": funcq, "shell": shellq, "primitive": primq},
        "build_tflops_fp64": total_flops / t64 / 1e12,
        "kernel_ms_fp64": kernel_ms,
        "class_roofline_fp64": class_roofline,
        "engines_ms": engines,
        "setup_s": setup_s,
        "roofline": {"bound": "fp64-valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
                     "traffic": measured_traffic(dom["ls"]),
                     "kernel": (f"k_eri_class<{','.join(map(str, dom['ls']))},double,JK>"
                                + (f" + k_eri_class<{','.join(map(str, dom['ls']))},double,JKFar>"
                                   if dom["primq_far"] else "")),
                     "kernel_ms": dom_ms,
                     "note": "ceiling: FP64 vector ALU peak (the Rys kernels issue no MFMA); "
                             "algorithmic flops = the fixed Rys flop count per primitive quartet "
                             "x the primitive quartets evaluated after screening; a one-node "
                             "class is two launches, timed together: near quartets, and far "
                             "ones in the point-charge form, which evaluates the same primitive "
                             "quartets with fewer executed flops",
                     "far_primitive_quartets": int(dom["primq_far"]),
                     "executed_flops_far_form": int(dom["primq_far"] * metrics.FAR_FLOPS.get(
                         tuple(dom["ls"]), 0))},
    }
    # full SCF (DIIS, tol 1e-8) in FP64 and in mixed precision on the same basis
    from segfault_py_amd import hf
    if not args.no_scf:
        # the process's first SCF pays one-time costs (code objects of the
        # one-electron and solver kernels loaded on first launch, GEMM and
        # eigensolver set-up): an untimed four-iteration SCF on a basis of its
        # own pays them, so the timed SCFs below (each still capturing its
        # own J/K graph) are steady-state
        import torch
        hf.scf_diis(mol, precision=0, maxiter=3)
        torch.cuda.synchronize()
        r64 = hf.scf_diis(mol, basis=basis, precision=0)
        rmx = hf.scf_diis(mol, basis=basis, precision=1)
        line["scf"] = {"energy_fp64": r64.energy, "energy_mixed": rmx.energy,
                       "dE_scf_mixed_vs_fp64": abs(rmx.energy - r64.energy),
                       "converged": bool(r64.converged and rmx.converged),
                       "iterations_fp64": r64.iterations, "iterations_mixed": rmx.iterations,
                       "wall_s_fp64": r64.wall_s, "wall_s_mixed": rmx.wall_s}
    if not args.no_scf:
        line["alkane_mixed"] = alkane_workload(args.steps)
        line["gradient"] = gradient_workload()
        line["published"] = published_workload()
        line["cc_pvtz"] = cc_pvtz_workload(args.steps)
    if not args.no_batch:
        del basis
        line["batch"] = batch_workload(args.batch_size)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(engine.Basis(mol).screen(tau, tau_mixed), mol, work,
                                            args.cpu_seconds)
    print(json.dumps(line))


def mixed_sweep(steps: int) -> dict:
    """(H2O)20: mixed J/K builds with the FP32 quartets' primitive threshold
    tau_prim swept, each against the FP64 build: time per build, and the
    two-electron energy error at the converged FP64 SCF density."""
    import torch
    from segfault_py_amd import engine, hf, workloads as W

    mol = W.config_molecule(1)
    b = engine.Basis(mol).screen(1e-12, 1e-5)
    r = hf.scf_diis(mol, basis=b, tol=1e-9)
    P = r.P.contiguous()
    J, K = torch.empty_like(P), torch.empty_like(P)

    def build(prec):
        b.set_tuning(b.autotune(P, prec)[0])
        b.jk(P, prec, out=(J, K))
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            b.jk(P, prec, out=(J, K))
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / steps * 1e3, float(0.5 * (P * (J - 0.5 * K)).sum())

    ms64, e64 = build(0)
    out = {"workload": f"(H2O)20, N={mol.nbf}, tau_mixed 1e-5, density of the converged FP64 SCF",
           "fp64_ms": ms64, "points": []}
    for tp in (-1.0, 1e-12, 1e-11, 1e-10, 1e-9):
        b.set_prim_tau(tp)
        ms, e = build(1)
        out["points"].append({"tau_prim_fp32": tp if tp >= 0 else 1e-13, "mixed_ms": ms,
                              "speedup": ms64 / ms, "dE2": abs(e - e64)})
    b.set_prim_tau(-1.0)
    return out


def published_workload() -> dict:
    """The reference's own benchmark (src/hf/benchmark/molecule_test.json,
    copied to tests/golden/): every molecule's cc-pVDZ (Cartesian) RHF by
    the DIIS SCF (tol 1e-8, Schwarz 1e-12), GPU wall time per molecule
    (handle setup, one-electron integrals, every J/K build, eigh and DIIS
    included) beside the published energy and wall time (the reference's
    scf_diis with its N^4 tensor, measured on 64 CPU cores,
    molecule_test.txt:1-3)."""
    import torch
    from segfault_py_amd import hf
    from segfault_py_amd.molecule import Molecule

    gold = os.path.join(ROOT, "tests", "golden")
    data = json.load(open(os.path.join(gold, "molecule_test.json")))
    basis = os.path.join(gold, "cc-pvdz-HCNO.dat")
    rows = []
    # one untimed run pays the process's first-call costs (solver set-up)
    hf.scf_diis(Molecule(os.path.join(gold, "molecules", "water.xyz"), basis))
    for group, mols in data.items():
        for name, ref in mols.items():
            mol = Molecule(os.path.join(gold, "molecules", f"{name}.xyz"), basis)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            r = hf.scf_diis(mol, tol=1e-8)
            torch.cuda.synchronize()
            wall = time.perf_counter() - t0
            rows.append({"molecule": name, "ncgf": mol.nbf, "gpu_wall_s": round(wall, 4),
                         "ref_wall_s": ref["wtime"], "speedup": round(ref["wtime"] / wall, 1),
                         "dE_vs_published": abs(r.energy - ref["etot"]),
                         "iterations": r.iterations, "converged": bool(r.converged)})
    gpu_total = sum(x["gpu_wall_s"] for x in rows)
    ref_total = sum(x["ref_wall_s"] for x in rows)
    return {"workload": f"{len(rows)} molecules of the reference's benchmark, cc-pVDZ "
                        "(Cartesian), DIIS SCF tol 1e-8, Schwarz 1e-12",
            "molecules": len(rows),
            "all_converged": all(x["converged"] for x in rows),
            "max_dE_vs_published": max(x["dE_vs_published"] for x in rows),
            "gpu_wall_s_total": gpu_total, "ref_wall_s_total": ref_total,
            "speedup_total": ref_total / gpu_total,
            "ref_hardware": "64 CPU cores (molecule_test.txt)",
            "largest": max(rows, key=lambda x: x["ncgf"]),
            "per_molecule": rows}


def cc_pvtz_workload(steps: int) -> dict:
    """Reference molecules in the reference's own cc-pVTZ basis
    (src/hf/data/basis/cc-pvtz.dat, copied to tests/golden/: f shells on C
    and O): FP64 J/K build time after autotuning, the share of it in the
    f-shell quartet classes (HIP events per class, one stream), and the DIIS
    SCF (tol 1e-8) wall time and energy.  No published number covers these."""
    import torch
    from segfault_py_amd import engine, hf
    from segfault_py_amd.molecule import Molecule

    gold = os.path.join(ROOT, "tests", "golden")
    basis = os.path.join(gold, "cc-pvtz-HCO.dat")
    rows = []
    for name in ("water", "ethanol", "1-hexanol"):
        mol = Molecule(os.path.join(gold, "molecules", f"{name}.xyz"), basis)
        b = engine.Basis(mol).screen(1e-12, 1e-5)
        nocc = int(mol.numbers.sum()) // 2
        P = torch.tensor(synthetic_density(mol.nbf, nocc), dtype=torch.float64, device="cuda")
        J, K = torch.empty_like(P), torch.empty_like(P)
        b.set_tuning(b.autotune(P, 0)[0])
        b.jk(P, 0, out=(J, K))
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            b.jk(P, 0, out=(J, K))
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t) / steps * 1e3
        _, _, prof = b.jk_profile(P, 0, out=(J, K))
        f_ms = float(sum(prof[2 * c] + prof[2 * c + 1]
                         for c in range(engine.NUM_CLASSES_SPD, engine.NUM_QUARTET_CLASSES)))
        t = time.perf_counter()
        r = hf.scf_diis(mol, basis=b, tol=1e-8)
        torch.cuda.synchronize()
        rows.append({"molecule": name, "ncgf": mol.nbf, "fp64_build_ms": round(ms, 3),
                     "f_class_share": round(f_ms / float(prof.sum()), 3),
                     "scf_wall_s": round(time.perf_counter() - t, 3), "energy": r.energy,
                     "iterations": r.iterations, "converged": bool(r.converged)})
    return {"workload": "reference molecules, cc-pVTZ (Cartesian; f on C and O), Schwarz 1e-12",
            "per_molecule": rows}


def alkane_workload(steps: int) -> dict:
    """Config 3: zigzag C16H34 (C 4s3p1d, H 3s1p, N = 508), quartets with
    Schwarz bound < 1e-5 in FP32: J/K builds/s FP64 vs mixed and the SCF
    energy error of mixed against FP64."""
    import torch
    from segfault_py_amd import engine, hf, workloads as W

    mol = W.config_molecule(2)
    b = engine.Basis(mol).screen(1e-12, 1e-5)
    nocc = int(mol.numbers.sum()) // 2
    P = torch.tensor(synthetic_density(mol.nbf, nocc), dtype=torch.float64, device="cuda")
    J, K = torch.empty_like(P), torch.empty_like(P)
    out = {"workload": f"C16H34, 50 atoms, N={mol.nbf}, C 4s3p1d / H 3s1p, tau_mixed 1e-5"}
    for name, prec in (("fp64", 0), ("mixed", 1)):
        b.set_tuning(b.autotune(P, prec)[0])
        b.jk(P, prec, out=(J, K))
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            b.jk(P, prec, out=(J, K))
        torch.cuda.synchronize()
        out[f"{name}_builds_per_s"] = steps / (time.perf_counter() - t)
        r = hf.scf_diis(mol, basis=b, precision=prec, tol=1e-9)
        out[f"energy_{name}"] = r.energy
        out[f"scf_wall_s_{name}"] = r.wall_s
    out["mixed_speedup"] = out["mixed_builds_per_s"] / out["fp64_builds_per_s"]
    out["dE_mixed_vs_fp64"] = abs(out["energy_mixed"] - out["energy_fp64"])
    c = b.counts()[0]
    out["fp32_fraction_of_shell_quartets"] = float(c[:, 2].sum() / max(c[:, 0].sum(), 1))
    return out


def gradient_workload() -> dict:
    """Config 4: ten waters (30 atoms, s/p/d basis): SCF + analytic RHF
    gradient (derivative-ERI kernels) all in FP64, and all in mixed precision
    (SCF and derivative quartets below tau_mixed in FP32), steady-state time
    of a gradient call (and the first call's, with its one-time setup); max
    gradient difference mixed vs FP64 (the finite-difference checks are in
    tests/test_gpu_grad.py)."""
    import torch
    from segfault_py_amd import engine, grad as G, hf, workloads as W

    mol = W.config_molecule(3)
    b = engine.Basis(mol).screen(1e-12, 1e-5)
    out = {"workload": f"(H2O)10, 30 atoms, N={mol.nbf}, s/p + d(O)"}
    g = {}
    for name, prec in (("fp64", 0), ("mixed", 1)):
        r = hf.scf_diis(mol, basis=b, precision=prec, tol=1e-10)
        # the first call also builds the reduction lists and buffers (once
        # per screened basis); the second is the steady-state cost
        for call in ("first", "steady"):
            torch.cuda.synchronize()
            t = time.perf_counter()
            g[name] = G.rhf_gradient(mol, r, b, prec)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t
            if call == "first" and name == "fp64":
                out["gradient_s_first_call"] = dt
        out[f"gradient_s_{name}"] = dt
        out[f"energy_{name}"] = r.energy
    out["max_dgrad_mixed_vs_fp64"] = float(np.max(np.abs(g["mixed"] - g["fp64"])))
    out["max_abs_gradient"] = float(np.max(np.abs(g["fp64"])))
    return out


def gradient_calls(ncall: int) -> dict:
    """(H2O)10 FP64 SCF, then ncall FP64 nuclear gradients (the first one
    includes its one-time setup); wall seconds of each call."""
    import torch
    from segfault_py_amd import engine, grad as G, hf, workloads as W

    mol = W.config_molecule(3)
    b = engine.Basis(mol).screen(1e-12, 1e-5)
    r = hf.scf_diis(mol, basis=b, precision=0, tol=1e-10)
    times = []
    for _ in range(ncall):
        torch.cuda.synchronize()
        t = time.perf_counter()
        G.rhf_gradient(mol, r, b, 0)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t)
    return {"workload": f"(H2O)10, 30 atoms, N={mol.nbf}, s/p + d(O)", "gradient_s": times}


def batch_workload(nmol: int) -> dict:
    """Config 5: nmol uneven molecules (water clusters of 1-8 waters and
    CkH2k+2, k = 1..7, s/p basis) solved to 1e-8 Eh by the batched DIIS SCF:
    one BatchBasis, autotuned persistent class kernels, J/K replayed from a
    hipGraph every iteration.  Throughput = molecules / wall time of the
    whole batch (setup included); latency = one molecule solved alone."""
    import torch
    from segfault_py_amd import engine, hf, workloads as W

    mols = W.batch(nmol, seed=5)
    # warm the batched eigh / solve library paths at the batch's shape (one-time
    # library initialisation, not part of the workload)
    nmax = max(m.nbf for m in mols)
    A = torch.randn(nmol, nmax, nmax, dtype=torch.float64, device="cuda")
    torch.linalg.eigh(A @ A.transpose(1, 2) + nmax * torch.eye(nmax, dtype=torch.float64, device="cuda"))
    torch.linalg.solve_ex(torch.eye(8, dtype=torch.float64, device="cuda").repeat(nmol, 1, 1),
                          torch.ones(nmol, 8, dtype=torch.float64, device="cuda"))
    del A
    out = {"workload": f"{nmol} molecules, 3-24 atoms (water clusters and alkanes), "
                       "even-tempered s/p basis, Schwarz 1e-12, SCF tol 1e-8",
           "atoms": int(sum(m.natom for m in mols)), "nbf_max": int(max(m.nbf for m in mols))}
    for name, prec in (("fp64", 0), ("mixed", 1)):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        bb = engine.BatchBasis(mols).screen(1e-12, 1e-5)
        res = hf.scf_batch(mols, precision=prec, basis=bb)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        c, fq, pq = bb.counts()
        out[name] = {"molecules_per_s": nmol / wall, "wall_s": wall, "scf_s": res.scf_s,
                     "jk_s": res.jk_s, "iterations": res.n_iter,
                     "jk_ms_per_iteration": 1e3 * res.jk_s / max(res.n_iter + 1, 1),
                     "converged": int(sum(res.converged)),
                     "shell_quartets_per_build": int(c[:, 0].sum()),
                     "function_quartets_per_build": int(fq),
                     "tuned_shapes": res.tuning[0][:, (0,) if prec == 0 else (1, 2), :]
                     .reshape(engine.NUM_QUARTET_CLASSES, -1).tolist() if res.tuning is not None else None}
        out[name]["_energies"] = res.energies
        del bb
    e64, emx = out["fp64"].pop("_energies"), out["mixed"].pop("_energies")
    out["max_dE_mixed_vs_fp64"] = max(abs(a - b) for a, b in zip(e64, emx))
    # latency: a few molecules solved alone (own handle, FP64)
    sample = sorted({0, nmol // 3, (2 * nmol) // 3, nmol - 1})
    lat, dE = [], []
    for k in sample:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = hf.scf_diis(mols[k], tol=1e-8)
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - t0)
        dE.append(abs(r.energy - e64[k]))
    out["single_molecule_latency_s"] = {"median": float(np.median(lat)), "max": float(max(lat)),
                                        "molecules": sample}
    out["max_dE_batch_vs_single"] = float(max(dE))
    out["sequential_molecules_per_s"] = 1.0 / float(np.median(lat))
    return out


if __name__ == "__main__":
    main()
