"""G0W0 solves/s on one MI355X: mixed-precision (BF16x3 Pi, FP64 factorisation) vs FP64.

Workload: one G0W0 contour-deformation solve at nocc 40 / nvir 400 / naux 1800 with 48
imaginary frequencies and Newton QP solves for HOMO-5..LUMO+5, on seeded synthetic inputs
(eps_occ ~ U[-1.5,-0.3], eps_vir ~ U[0.05,3.0], B = U diag(0.5 e^{-k/20}) V^T), resident in
HBM before timing.  A step is one full solve (Pi at 48 frequencies -> Cholesky -> L^-1 ->
Wc_nm -> Newton).  Prints one JSON line.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu-baseline]
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

NOCC, NVIR, NAUX, NF = 40, 400, 1800, 48
METRIC = "G₀W₀ solves/s, nocc40/nvir400/naux1800: mixed vs FP64 GPU at ≤1 meV"
PEAK = {"bf16": 2.5e15, "fp32": 157.3e12, "fp64": 78.6e12}
MEV = 1.0 / 27211.386245988


PI_KERNELS = {0: "pi_syrk_bf16x3_kernel<32>", 1: "pi_syrk_bf16x3_kernel<64>",
              2: "pi_syrk_bf16x3_wide_kernel", 3: "pi_syrk_bf16x3_db_kernel<32>",
              4: "pi_syrk_bf16x3_db_kernel<64>", 5: "pi_syrk_bf16x3_m32_kernel<64>",
              6: "pi_syrk_bf16x3_m32_kernel<32>", 7: "pi_syrk_bf16x3_ws_kernel<64, 4>",
              8: "pi_syrk_bf16x3_ws2_kernel", 9: "pi_syrk_bf16x3_ws_kernel<32, 4>",
              10: "pi_syrk_bf16x3_ws_kernel<64, 8>"}
PI_KERNELS_F64 = {0: "pi_syrk_kernel<double, 16>", 1: "pi_f64_pair_kernel<16, false>",
                  2: "pi_f64_pair_kernel<32, false>", 3: "pi_f64_pair_kernel<32, true>",
                  4: "pi_f64_pair_kernel<16, true>"}
TRAFFIC = os.path.join(ROOT, "profiles", "pi_hbm_traffic.json")


def measured_traffic(kernel):
    """HBM bytes per launch of ``kernel`` at this workload, from the rocprofv3 FETCH_SIZE /
    WRITE_SIZE passes recorded in profiles/ (scaled by the copy calibration there); None when
    no record matches the kernel the autotuner picked."""
    try:
        rec = json.load(open(TRAFFIC))
    except (OSError, ValueError):
        return None
    k = rec.get("kernels", {}).get(kernel)
    return k["bytes_per_launch"] if k else None


COUNTERS = os.path.join(ROOT, "profiles", "mfma_counters.json")


def measured_mfma(kernel):
    """MFMA-pipe busy fraction and issued / algorithmic MFMA flops of ``kernel`` from the
    rocprofv3 counter passes recorded in profiles/mfma_counters.json ({} if absent)."""
    try:
        rec = json.load(open(COUNTERS)).get("kernels", {}).get(kernel)
    except (OSError, ValueError):
        return {}
    if not rec:
        return {}
    return {k: rec[k] for k in ("mfma_busy", "mfma_issued_over_algorithmic") if k in rec}


def pi_flops(nov, naux, nf):
    """Algorithmic flops of the Pi launch: the symmetric rank-N_ov update of every
    frequency, N_ov * naux * (naux + 1) per frequency (2 flops per multiply-add over the
    naux (naux+1)/2 distinct entries)."""
    return float(nf) * nov * naux * (naux + 1)


def solve_flops(nov, naux, nf, k):
    """Algorithmic flops of one solve: Pi + Cholesky & inverse (naux^3) + Wc_nm (2 naux^2 k
    + 2 naux k) per frequency."""
    return nf * (nov * naux * (naux + 1) + naux ** 3 + 2 * naux * naux * k + 2 * naux * k)


def host_description():
    """CPU model, logical CPUs, the BLAS threads numpy will use and numpy's BLAS/LAPACK build."""
    model = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    threads, blas = None, {}
    try:
        import threadpoolctl
        for p in threadpoolctl.threadpool_info():
            if p.get("user_api") == "blas":
                threads = p.get("num_threads")
                blas = {"internal_api": p.get("internal_api"), "version": p.get("version"),
                        "architecture": p.get("architecture")}
                break
    except Exception:
        pass
    try:
        cfg = np.show_config(mode="dicts")["Build Dependencies"]
        blas["numpy_blas"] = cfg["blas"].get("name") + " " + str(cfg["blas"].get("version"))
        blas["numpy_lapack"] = cfg["lapack"].get("name") + " " + str(cfg["lapack"].get("version"))
    except Exception:
        pass
    return {"cpu_model": model, "logical_cpus": os.cpu_count(), "blas_threads": threads,
            "blas": blas}


def cpu_baseline(e, B, nocc, targets):
    """The numpy/scipy FP64 oracle (restatement of the reference CPU path: per-frequency
    scaled DGEMM, Cholesky inverse, per-target contraction, CD self-energy and Newton) run
    once in full on the same workload: all 48 frequencies, every target."""
    from oracle import gw_oracle as O
    fr, wt = O.get_scaled_legendre_roots(NF, 0.5)
    host = host_description()
    t0 = time.perf_counter()
    O.g0w0_cd(e, B, nocc, np.asarray(targets), fr, wt)
    t_solve = time.perf_counter() - t0
    return {"value": 1.0 / t_solve, "unit": "solves/s",
            "cores": host["blas_threads"] or host["logical_cpus"], "kind": "port",
            "sample": f"one full solve ({NF} frequencies, {len(targets)} targets) of the "
                      f"numpy/scipy FP64 oracle in {t_solve:.2f} s",
            "host": host}


def run_batch(args):
    """Batched workload: 256 instances, nocc ~ U{8..60}, nvir = 9 nocc, naux = 4 (nocc+nvir),
    32 frequencies, HOMO and LUMO of each; one step = one qx_g0w0_batch_ov_dev call (a
    captured-graph replay after the first two calls)."""
    import torch
    from quasix_amd import gw, synthetic
    nf = 32
    shapes = synthetic.batch_shapes(args.batch_size, seed=2)
    dev = torch.device("cuda", 0)
    inst = []
    for q, (o, v, a) in enumerate(shapes):
        e, B = synthetic.make_inputs_device(o, v, a, seed=1000 + q, device=dev)
        inst.append((e, B, o, np.array([o - 1, o], np.int32)))
        if q % 32 == 31:
            print(f"generated {q + 1} instances", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    flops = sum(nf * (o * v) * a * (a + 1) for o, v, a in shapes)

    def timed(prec, steps, warmup):
        s = gw.G0W0BatchSolver(inst, nfreq=nf, precision=prec)
        s.set_profiling(True)
        s.run()                                   # direct run: tunes, records phases
        prof = s.profile()
        s.set_profiling(False)
        for _ in range(max(warmup - 1, 1)):
            s.run()                               # captures the graph
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            s.run()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / steps
        print(f"{prec}: {dt * 1e3:.1f} ms per batch", file=sys.stderr, flush=True)
        res = s.results()
        st = dict(waves=s.stats.waves, graph_replay=s.stats.graph_replay,
                  resolved=s.stats.resolved_instances, pi_variant=s.stats.pi_variant)
        del s
        return res, dt, prof, st

    r_mix, dt_mix, prof, st_mix = timed("bf16x3", args.steps, args.warmup)
    r_64, dt_64, prof64, st_64 = timed("fp64", max(1, args.steps // 3), 2)
    dE = max(float(np.max(np.abs(a["qp_energies"] - b["qp_energies"])))
             for a, b in zip(r_mix, r_64))
    ms_pi, n_pi = prof["pi"]
    line = {
        "metric": f"batched G0W0 instances/s ({len(shapes)} instances, one call)",
        "value": len(shapes) / dt_mix, "unit": "instances/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": dt_mix * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16x3+f64", "data": "synthetic (device-generated, same distributions)",
        "config": {"workload": f"{len(shapes)} instances nocc~U{{8..60}}, nvir=9 nocc, "
                               "naux=4(nocc+nvir), 32 frequencies, HOMO+LUMO",
                   "instances": len(shapes), "nfreq": nf},
        "fp64_instances_per_s": len(shapes) / dt_64, "speedup_vs_fp64": dt_64 / dt_mix,
        "max_abs_dE_qp_meV": dE / MEV, "stats_mixed": st_mix, "stats_fp64": st_64,
        "phase_ms_mixed_direct_run": {k: round(v[0], 3) for k, v in prof.items() if v[1]},
        "phase_ms_fp64_direct_run": {k: round(v[0], 3) for k, v in prof64.items() if v[1]},
        "pi_flops": flops, "pi_tflops_mixed": flops / (ms_pi * 1e-3) / 1e12,
    }
    print(json.dumps(line), flush=True)


def run_evgw(args):
    """evGW workload: nocc 100 / nvir 1200 / naux 5000 (B 4.8 GB FP64, resident across the
    iterations), 64 frequencies, 8 eigenvalue iterations updating HOMO-5..LUMO+5; the mixed
    path (BF16x3 Pi) against the FP64 GPU path; reports the converged HOMO-LUMO gap error."""
    import torch
    from quasix_amd import gw, synthetic
    nocc, nvir, naux, nf, nit = 100, 1200, 5000, 64, 8
    dev = torch.device("cuda", 0)
    e, B = synthetic.make_inputs_device(nocc, nvir, naux, seed=4, device=dev)
    torch.cuda.synchronize()
    print("inputs ready", file=sys.stderr, flush=True)
    tg = gw.frontier_targets(nocc, nocc + nvir)
    out = {}
    for prec in ("bf16x3", "fp64"):
        t0 = time.perf_counter()
        r = gw.evgw_core(e, B, nocc, tg, n_iter=nit, damping=0.5, conv_tol=0.0, nfreq=nf,
                         precision=prec, progress=lambda it, dm: print(
                             f"{prec} iteration {it}: max |dE| {dm:.3e} Ha", file=sys.stderr,
                             flush=True))
        dt = time.perf_counter() - t0
        qp = r["qp_energies"]
        out[prec] = dict(seconds=dt, gap=float(qp[nocc] - qp[nocc - 1]), qp=qp[tg])
    dgap = abs(out["bf16x3"]["gap"] - out["fp64"]["gap"])
    line = {
        "metric": "evGW iterations/s (nocc100/nvir1200/naux5000, 64 frequencies, 8 iterations)",
        "value": nit / out["bf16x3"]["seconds"], "unit": "iterations/s", "n_gpus": 1,
        "steps": nit, "warmup": 0, "ms_per_step": out["bf16x3"]["seconds"] / nit * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16x3+f64", "data": "synthetic (device-generated, same distributions)",
        "config": {"workload": "evgw_cd nocc100/nvir1200/naux5000, 64 frequencies, 8 iterations, "
                               "HOMO-5..LUMO+5", "nocc": nocc, "nvir": nvir, "naux": naux,
                   "nfreq": nf, "iterations": nit},
        "fp64_iterations_per_s": nit / out["fp64"]["seconds"],
        "speedup_vs_fp64": out["fp64"]["seconds"] / out["bf16x3"]["seconds"],
        "gap_eV_fp64": out["fp64"]["gap"] * 27.211386245988,
        "gap_error_meV": dgap / MEV,
        "max_abs_dE_qp_meV": float(np.max(np.abs(out["bf16x3"]["qp"] - out["fp64"]["qp"]))) / MEV,
    }
    print(json.dumps(line), flush=True)


def run_ac(args):
    """AC self-energy workload (the reference's live method, correlation_ac.rs:241-449):
    Sigma_c(i nu) of HOMO-5..LUMO+5 at nocc 40 / nvir 400 / naux 1800 on the grid of 48 scaled
    Gauss-Legendre frequencies (nu = 0 and the nodes below 5 Ha), FP64 and mixed; one step is
    one qx_sigma_ac_dev call (Pi, factorisation, Wc_nm, quadrature) on resident inputs."""
    import torch
    from quasix_amd import gw, synthetic
    e, B = synthetic.make_inputs(NOCC, NVIR, NAUX, seed=1)
    dev = torch.device("cuda", 0)
    e_d, B_d = torch.as_tensor(e, device=dev), torch.as_tensor(B, device=dev)
    tg = gw.frontier_targets(NOCC, NOCC + NVIR)
    pairs = gw.ov_pairs(B_d, NOCC, NVIR, tg)
    out = {}
    for prec in ("bf16x3", "fp64"):
        for _ in range(args.warmup):
            sig, _ = gw.sigma_ac(e_d, B_d, NOCC, pairs, tg, nw=NF, precision=prec)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            sig, _ = gw.sigma_ac(e_d, B_d, NOCC, pairs, tg, nw=NF, precision=prec)
        torch.cuda.synchronize()
        out[prec] = ((time.perf_counter() - t0) / args.steps, sig)
    d = float(np.max(np.abs(out["bf16x3"][1] - out["fp64"][1])))
    line = {
        "metric": "AC Sigma_c(i nu) evaluations/s (nocc40/nvir400/naux1800, 48 frequencies)",
        "value": 1.0 / out["bf16x3"][0], "unit": "evaluations/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": out["bf16x3"][0] * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16x3+f64", "data": "synthetic (seeded U/V QR; inputs resident in HBM)",
        "config": {"workload": "sigma_ac nocc40/nvir400/naux1800, 48 frequencies, "
                               "HOMO-5..LUMO+5", "nfreq": NF, "targets": len(tg)},
        "fp64_ms_per_step": out["fp64"][0] * 1e3,
        "max_abs_dsigma_meV": d / MEV,
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--precision", default="bf16x3")
    ap.add_argument("--mixed-only", action="store_true",
                    help="skip the FP64 comparison (short profiled runs)")
    ap.add_argument("--workload", choices=["single", "batch", "evgw", "ac"], default="single",
                    help="batch: the 256-instance grouped workload; evgw: 8 evGW iterations at "
                         "nocc 100 / nvir 1200 / naux 5000; ac: the AC self-energy on the "
                         "imaginary axis at the headline size (each its own JSON line)")
    ap.add_argument("--batch-size", type=int, default=256)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 1 GiB device-to-device 3 times (scales HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the path shards "
              "over frequencies/instances across GPUs only by design, see the design notes)",
              file=sys.stderr)
        sys.exit(2)

    import torch
    if args.copy_calibration:
        n = 1 << 27                                   # 2^27 doubles = 1 GiB
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": n * 8, "reps": 3}), flush=True)
        return
    if args.workload == "batch":
        run_batch(args)
        return
    if args.workload == "evgw":
        run_evgw(args)
        return
    if args.workload == "ac":
        run_ac(args)
        return
    from quasix_amd import gw, synthetic

    e, B = synthetic.make_inputs(NOCC, NVIR, NAUX, seed=1)
    targets = gw.frontier_targets(NOCC, NOCC + NVIR)
    dev = torch.device("cuda", 0)
    e_d = torch.as_tensor(e, device=dev)
    B_d = torch.as_tensor(B, device=dev)
    nov = NOCC * NVIR
    k = sum(NVIR if n < NOCC else NOCC for n in targets)

    def timed(prec, steps, warmup, profile=False):
        s = gw.G0W0Solver(e_d, B_d, NOCC, targets, nfreq=NF, precision=prec)
        for _ in range(warmup):
            s.run()
        s.set_profiling(profile)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            s.run()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / steps
        prof = s.profile() if profile else None
        s.set_profiling(False)
        res = s.result()
        res["pi_variant"] = s.stats.pi_variant
        return res, dt, prof

    r_mix, dt_mix, prof = timed(args.precision, args.steps, args.warmup, profile=True)
    assert np.all(r_mix["state"] == 0)
    if args.mixed_only:
        r_64, dt_64, prof64 = r_mix, float("nan"), {}
    else:
        r_64, dt_64, prof64 = timed("fp64", max(2, args.steps // 4), 1, profile=True)
        assert np.all(r_64["state"] == 0)
    dE = float(np.max(np.abs(r_mix["qp_energies"] - r_64["qp_energies"])))

    ms_pi, n_pi = prof["pi"]
    t_pi = ms_pi / max(n_pi, 1) * 1e-3
    f_pi = pi_flops(nov, NAUX, NF)
    peak = PEAK["bf16"] if args.precision == "bf16x3" else PEAK["fp32" if args.precision == "fp32" else "fp64"]
    achieved = f_pi / t_pi
    phases = {k2: round(v[0] / max(v[1], 1), 4) for k2, v in prof.items() if v[1]}
    phases64 = {k2: round(v[0] / max(v[1], 1), 4) for k2, v in prof64.items() if v[1]}

    kernel = (PI_KERNELS[r_mix["pi_variant"]] if args.precision == "bf16x3"
              else "pi_syrk_kernel<float, 32>" if args.precision == "fp32"
              else PI_KERNELS_F64[r_mix["pi_variant"]])
    fp64_pi = None
    if prof64.get("pi", (0, 0))[1]:
        ms_pi64, n_pi64 = prof64["pi"]
        t_pi64 = ms_pi64 / n_pi64 * 1e-3
        kernel64 = PI_KERNELS_F64[r_64["pi_variant"]]
        fp64_pi = {"kernel": kernel64, "bound": "mfma", "launch_ms": t_pi64 * 1e3,
                   "achieved": f_pi / t_pi64 / 1e12, "peak": PEAK["fp64"] / 1e12,
                   "unit": "TFLOP/s", "frac": f_pi / t_pi64 / PEAK["fp64"],
                   "traffic": measured_traffic(kernel64), **measured_mfma(kernel64)}
    line = {
        "metric": METRIC,
        "value": 1.0 / dt_mix,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt_mix * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16x3+f64" if args.precision == "bf16x3" else args.precision,
        "data": "synthetic (seeded U/V QR, s_k = 0.5 exp(-k/20); inputs resident in HBM)",
        "config": {"workload": "g0w0_cd nocc40/nvir400/naux1800, 48 imaginary frequencies, "
                               "Newton QP for HOMO-5..LUMO+5",
                   "nocc": NOCC, "nvir": NVIR, "naux": NAUX, "nfreq": NF,
                   "targets": len(targets), "pairs_k": k, "pi_precision": args.precision,
                   "factorisation": "fp64"},
        "fp64_solves_per_s": 1.0 / dt_64,
        "fp64_ms_per_step": dt_64 * 1e3,
        "speedup_vs_fp64": dt_64 / dt_mix,
        "max_abs_dE_qp_meV": dE / MEV,
        "phase_ms_mixed": phases,
        "phase_ms_fp64": phases64,
        "solve_tflops_mixed": solve_flops(nov, NAUX, NF, k) / dt_mix / 1e12,
        "roofline": {"kernel": kernel, "bound": "mfma", "achieved": achieved / 1e12,
                     "peak": peak / 1e12, "unit": "TFLOP/s", "frac": achieved / peak,
                     "traffic": measured_traffic(kernel), "flops_per_launch": f_pi,
                     "launch_ms": t_pi * 1e3,
                     "traffic_source": "rocprofv3 FETCH_SIZE/WRITE_SIZE passes of this "
                                       "kernel recorded in profiles/pi_hbm_traffic.json "
                                       "(not measured in this run)",
                     **measured_mfma(kernel)},
        "fp64_pi_roofline": fp64_pi,
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(e, B, NOCC, targets)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
