"""Benchmark: Π point-scales/sec of the filter + Π pipeline on one MI355X.

Workload (one step = one compute_Π! over the whole field at one scale): spherical 0.25° lon-lat
grid (1440×720 cell centres), FP32 velocity from a random-phase spherical-harmonic streamfunction
(E(l) ∝ l⁻³, 8 ≤ l ≤ 400, RMS 0.2 m/s), 30 % land from a smooth Gaussian random field, ZeroFill,
Gaussian α = 4 on haversine distance, ℓ = 200 km. Inputs are resident in HBM before timing.

The measured engine is the library's default for this grid (the ring engine: the reference's banded
footprint evaluated in Fourier space along the periodic longitude, FP64 inside). Also reported on the
same inputs: the fused FP32 direct-sum engine, the plain FP64 GPU direct sum (one thread per point,
CSR in HBM), the FP64 CPU path (the oracle, an OpenMP restatement of the reference's threaded
row-parallel gather) on this host's cores, and the rel-L2 / max error of Π against that oracle.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "coarsegrainingenergyfluxes.jl_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0      # MI355X HBM3E
FP32_PEAK_TFLOPS = 157.3   # MI355X dense FP32
SCALE = 200e3


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def timed_steps(step, steps, warmup, plan=None):
    import torch
    import cgef
    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    L = cgef._lib.load()
    if plan is not None:
        cgef._lib.check(L.cgef_profile_enable(plan.h, 1))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    kern_ms = None
    if plan is not None:
        kms, nl = C.c_double(), C.c_int64()
        cgef._lib.check(L.cgef_profile_read(plan.h, C.byref(kms), C.byref(nl)))
        cgef._lib.check(L.cgef_profile_enable(plan.h, 0))
        kern_ms = kms.value / max(nl.value, 1)
    return el, kern_ms


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--engine", default="auto", help="auto (ring) | banded | baseline")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-extra", action="store_true", help="skip the secondary GPU engines")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; --gpus above 1 is not supported (the multi-GPU "
              "decomposition is described, not run).", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import cgef
    from cgef import synthetic as S

    dev = "cuda:0"
    torch.cuda.set_device(0)
    log("generating the 0.25° spherical workload")
    u, v, lam, phi = S.sphere_turbulence()
    mask = S.sphere_land_mask()
    grid = cgef.StructuredGrid(cgef.SphericalGeometry(6.371e6), lam, phi, mask=mask)
    ker = cgef.GaussianKernel(4.0)
    npts = grid.npoints
    u32 = torch.tensor(u, dtype=torch.float32, device=dev)
    v32 = torch.tensor(v, dtype=torch.float32, device=dev)
    Pi = torch.empty_like(u32)
    ws = cgef.PiWorkspace(grid, (), torch.float32)
    t0 = time.perf_counter()
    plan = cgef.plan_filter(grid, ker, SCALE, engine=args.engine)
    torch.cuda.synchronize()
    info = plan.info
    log(f"plan ({plan.engine}) built in {time.perf_counter() - t0:.2f} s: {info.nentries} footprint entries, "
        f"mean taps {info.mean_taps:.1f}, max {info.max_taps}")

    def step():
        cgef.compute_pi_(Pi, u32, v32, None, grid, ker, SCALE, workspace=ws, filter_plan=plan)

    elapsed, kern_ms = timed_steps(step, args.steps, args.warmup, plan)
    ms_step = elapsed * 1e3 / args.steps
    value = npts * args.steps / elapsed
    _, E_gpu = cgef.compute_pi_(Pi, u32, v32, None, grid, ker, SCALE, workspace=ws, filter_plan=plan,
                                return_energy=True)
    torch.cuda.synchronize()
    Pi_gpu = Pi.detach().cpu().double().numpy()
    log(f"{plan.engine}: {ms_step:.3f} ms/step, dominant kernel {kern_ms:.3f} ms")

    if plan.engine == "ring":
        # k_contract: Ŵ[m][offset][band] once, forward spectra once, contracted spectra written once
        nf = 10
        mh = grid.shape[0] // 2 + 1
        nrows = grid.shape[1]
        noff = 2 * int(np.ceil(cgef.kernel_radius(ker, SCALE) / (6.371e6 * (phi[1] - phi[0])))) + 1
        traffic_alg = mh * noff * nrows * 8 + 2 * nf * mh * nrows * 16
        achieved = traffic_alg / (kern_ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": "k_contract", "kernel_ms": kern_ms,
                "algorithmic_bytes_per_launch": traffic_alg}
    else:
        flops = 2.0 * 9.0 * info.active_taps
        achieved = flops / (kern_ms * 1e-3) / 1e12
        roof = {"bound": "mfma", "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / FP32_PEAK_TFLOPS, "traffic": None, "kernel": "k_pi_filter", "kernel_ms": kern_ms,
                "algorithmic_flops_per_launch": flops}

    extra = {}
    if not args.no_extra:
        if plan.engine != "banded":
            log("fused FP32 direct-sum engine")
            bplan = cgef.plan_filter(grid, ker, SCALE, engine="banded")
            P2 = torch.empty_like(u32)
            el, km = timed_steps(lambda: cgef.compute_pi_(P2, u32, v32, None, grid, ker, SCALE, workspace=ws,
                                                          filter_plan=bplan), max(args.steps // 2, 3), 2, bplan)
            fl = 2.0 * 9.0 * bplan.info.active_taps
            extra["gpu_fp32_fused_direct_sum"] = {
                "value": npts * max(args.steps // 2, 3) / el, "unit": "point-scales/s",
                "ms_per_step": el * 1e3 / max(args.steps // 2, 3), "kernel": "k_pi_filter", "kernel_ms": km,
                "algorithmic_tflops": fl / (km * 1e-3) / 1e12}
            del bplan, P2
        log("FP64 GPU plain direct sum (one thread per point)")
        bplan = cgef.plan_filter(grid, ker, SCALE, engine="baseline")
        u64, v64 = u32.double(), v32.double()
        P64 = torch.empty_like(u64)
        el, _ = timed_steps(lambda: cgef.compute_pi_(P64, u64, v64, None, grid, ker, SCALE, filter_plan=bplan), 2, 1)
        tb = el / 2
        extra["gpu_fp64_plain_direct_sum"] = {"value": npts / tb, "unit": "point-scales/s", "ms_per_step": tb * 1e3,
                                              "speedup_of_measured_engine": value / (npts / tb)}
        del bplan, P64, u64, v64

    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        ncores = min(len(os.sched_getaffinity(0)), 16)
        log(f"FP64 CPU oracle on {ncores} threads (full field, one scale)")
        og = O.Grid([lam, phi], spherical=True, R=6.371e6, periodic=[True, False], period=[2 * np.pi, 0.0],
                    mask=mask, uniform=[True, True])
        uin, vin = u32.cpu().double().numpy(), v32.cpu().double().numpy()
        fp = O.Footprint(og, O.GAUSSIAN, 4.0, SCALE)
        t2 = time.perf_counter()
        ref, ub, vb, _ = O.compute_pi(og, uin, vin, kind=O.GAUSSIAN, alpha=4.0, ell=SCALE, fp=fp, nthreads=ncores)
        tc = time.perf_counter() - t2
        cpu = {"value": npts / tc, "unit": "point-scales/s", "cores": ncores, "kind": "port",
               "sample": f"full 1440x720 field at one scale (1 step), {tc:.1f} s"}
        diff = Pi_gpu - ref
        extra["rel_l2_vs_fp64_oracle"] = float(np.sqrt(np.sum(diff ** 2) / np.sum(ref ** 2)))
        extra["max_abs_err"] = float(np.max(np.abs(diff)))
        extra["rel_max_err"] = float(np.max(np.abs(diff)) / np.max(np.abs(ref)))
        Eo = O.energy(og, ub, vb)
        extra["energy_rel_err"] = float(abs(E_gpu - Eo) / abs(Eo))
        extra["speedup_vs_cpu"] = value / cpu["value"]
        log(f"parity: rel-L2 {extra['rel_l2_vs_fp64_oracle']:.3e}, CPU {tc:.2f} s")

    out = {
        "metric": "Π point-scales/sec at rel L2 ≤1e-5 vs FP64 oracle; % of HBM/MFMA roofline",
        "value": value,
        "unit": "point-scales/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if plan.engine == "banded" else "f32 in/out, f64 inside",
        "data": "synthetic",
        "config": {"workload": "spherical 0.25deg 1440x720 FP32, Gaussian alpha=4 haversine, l=200km, 30% land "
                               "ZeroFill, 1 scale, 1 snapshot", "engine": plan.engine,
                   "points": npts, "scales": 1, "snapshots": 1, "parallelism": "1 GPU"},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    out.update(extra)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
