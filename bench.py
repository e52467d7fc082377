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
COPY_CAL_BYTES = 512 << 20
PROFILE_SUMMARY = "ring_sphere1440x720_l200km.md"


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` on this workload, as the committed hardware-counter summary
    states them (FETCH_SIZE and WRITE_SIZE runs scaled by a copy kernel of known traffic), or None."""
    path = os.path.join(ROOT, "profiles", PROFILE_SUMMARY)
    try:
        with open(path) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 3 and parts[0] == "traffic_bytes_per_launch" and parts[1] == kernel:
                    return int(parts[2])
    except OSError:
        pass
    return None


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


def batched_leg(args, grid, mask, lam, phi):
    """Batched filtering spectrum: 8 seeded snapshots × 24 scales log-spaced 50–2000 km on the same
    grid and mask, Π (1440,720,24,8) + E(ℓ) + Ẽ(k_ℓ) per snapshot in one coarse_grain_batch! call
    (one forward transform shared by all 24 scales). Spot parity against the oracle at two scales."""
    import numpy as np
    import torch

    import cgef
    from cgef import synthetic as S

    dev = "cuda:0"
    log("batched leg: generating 8 snapshots")
    us, vs, _, _ = S.sphere_turbulence_batch(range(100, 108))
    U = torch.tensor(np.stack(us), dtype=torch.float32, device=dev)
    V = torch.tensor(np.stack(vs), dtype=torch.float32, device=dev)
    scales = np.geomspace(50e3, 2000e3, 24)
    ker = cgef.GaussianKernel(4.0)
    t0 = time.perf_counter()
    plans = [cgef.plan_filter(grid, ker, float(s)) for s in scales]
    torch.cuda.synchronize()
    tplan = time.perf_counter() - t0
    log(f"batched leg: 24 plans ({plans[0].engine}) built in {tplan:.1f} s")
    batch = cgef.CoarseGrainBatchResult.allocate(grid, len(scales), (8,), torch.float32, dev)
    ws = cgef.PiWorkspace(grid, (8,), torch.float32)

    def call():
        cgef.coarse_grain_batch_(batch, U, V, None, grid, scales=scales, kernel=ker, filter_plans=plans,
                                 workspaces=[ws])

    # warm-up: the first sweep runs eagerly (and measures the contraction variants), the second is
    # captured as one HIP graph; the timed sweeps replay it
    nrep = 3
    el, kern = timed_steps(call, nrep, 2)
    ms = el * 1e3 / nrep
    value = grid.npoints * len(scales) * 8 * nrep / el
    log(f"batched leg: {ms:.2f} ms per 8x24 call -> {value:.3e} point-scales/s")
    out = {"value": value, "unit": "point-scales/s", "ms_per_call": ms, "snapshots": 8, "scales": 24,
           "plan_build_s": tplan, "engine": plans[0].engine,
           "launch": "one HIP graph per sweep (captured on the second call, replayed)"}
    if not args.no_cpu:
        from oracle import oracle as O
        og = O.Grid([lam, phi], spherical=True, R=6.371e6, periodic=[True, False], period=[2 * np.pi, 0.0],
                    mask=mask, uniform=[True, True])
        ncores = min(len(os.sched_getaffinity(0)), 16)
        checks = {}
        for si in (0, 8):
            ref, ub, vb, _ = O.compute_pi(og, U[0].cpu().double().numpy(), V[0].cpu().double().numpy(),
                                          kind=O.GAUSSIAN, alpha=4.0, ell=float(scales[si]), nthreads=ncores)
            got = batch.Pi[0, si].cpu().double().numpy()
            checks[f"scale_{scales[si] / 1e3:.0f}km"] = {
                "rel_l2_pi": float(np.sqrt(np.sum((got - ref) ** 2) / np.sum(ref ** 2))),
                "energy_rel_err": float(abs(batch.cumulative_energy[0, si] - O.energy(og, ub, vb)) / O.energy(og, ub, vb))}
        out["parity_snapshot0"] = checks
        log(f"batched leg parity: {checks}")
    del plans, batch, ws, U, V
    torch.cuda.empty_cache()
    return out


def tophat_leg(args, grid, lam, phi, mask, u32, v32):
    """The pipeline's default kernel (the top-hat, Pipeline.jl:344) on the same 0.25° sphere and
    fields at ℓ = 200 km: the library's default engine for it is the reference's prefix-sum engine
    (per-row prefix sums); the ring engine evaluating the same cap is timed beside it."""
    import numpy as np
    import torch

    import cgef

    ker = cgef.TopHatKernel()
    Pi = torch.empty_like(u32)
    ws = cgef.PiWorkspace(grid, (), torch.float32)
    out = {"workload": "spherical 0.25deg 1440x720 FP32, top-hat l=200km (cap radius 100 km), 30% land ZeroFill"}
    for eng in ("auto", "ring"):
        plan = cgef.plan_filter(grid, ker, SCALE, engine=eng)
        steps = 20

        def step():
            cgef.compute_pi_(Pi, u32, v32, None, grid, ker, SCALE, workspace=ws, filter_plan=plan)

        el, _ = timed_steps(step, steps, 3)
        out[plan.engine] = {"ms_per_step": el * 1e3 / steps, "value": grid.npoints * steps / el,
                            "unit": "point-scales/s"}
        got = Pi.cpu().double().numpy()
        log(f"top-hat sphere {plan.engine}: {el * 1e3 / steps:.3f} ms/step")
        if not args.no_cpu:
            from oracle import oracle as O
            og = O.Grid([lam, phi], spherical=True, R=6.371e6, periodic=[True, False], period=[2 * np.pi, 0.0],
                        mask=mask, uniform=[True, True])
            oeng = O.ENGINE_PREFIX if plan.engine == "prefixsum" else O.ENGINE_BANDED
            ncores = min(len(os.sched_getaffinity(0)), 16)
            t0 = time.perf_counter()
            ref = O.compute_pi(og, u32.cpu().double().numpy(), v32.cpu().double().numpy(), kind=O.TOPHAT,
                               alpha=0.0, ell=SCALE, engine=oeng, nthreads=ncores)[0]
            out[plan.engine]["cpu_s"] = time.perf_counter() - t0
            out[plan.engine]["rel_l2_vs_fp64_oracle"] = float(np.sqrt(np.sum((got - ref) ** 2) / np.sum(ref ** 2)))
    return out


def cartesian_leg(args):
    """Plain method at small size: 256×256 doubly periodic Cartesian, L = 1000 km, FP64, no mask,
    Gaussian α = 6, ℓ = 8Δx, ZeroFill, Π with ρ₀ = 1025. The library's default engine here is the
    reference's (the separable Gaussian); the disk-footprint engines are timed beside it. Parity of
    each against the oracle's matching engine."""
    import numpy as np
    import torch

    import cgef
    from cgef import synthetic as S

    dev = "cuda:0"
    N, L = 256, 1.0e6
    dx = L / N
    ell = 8 * dx
    u, v, xs = S.cartesian_turbulence(N=N, L=L, seed=0)
    grid = cgef.StructuredGrid(cgef.CartesianGeometry(), xs, xs, periodic=(True, True))
    ker = cgef.GaussianKernel(6.0)
    U, V = torch.tensor(u, device=dev), torch.tensor(v, device=dev)
    Pi = torch.empty_like(U)
    ws = cgef.PiWorkspace(grid, (), torch.float64)
    out = {"workload": "cartesian 256x256 doubly periodic FP64, Gaussian alpha=6, l=8dx, rho0=1025"}
    refs = {}
    if not args.no_cpu:
        from oracle import oracle as O
        og = O.Grid([xs, xs], periodic=[True, True], period=[L, L], uniform=[True, True])
        ncores = min(len(os.sched_getaffinity(0)), 16)
        for eng, oeng in (("separable", O.ENGINE_SEPARABLE), ("banded", O.ENGINE_BANDED)):
            t0 = time.perf_counter()
            refs[eng] = 1025.0 * O.compute_pi(og, u, v, kind=O.GAUSSIAN, alpha=6.0, ell=ell, engine=oeng,
                                              nthreads=ncores)[0]
            out[f"cpu_{eng}_s"] = time.perf_counter() - t0
        refs["ring"] = refs["banded"]
        out["cpu_cores"] = ncores
    for eng in ("auto", "banded", "ring"):
        plan = cgef.plan_filter(grid, ker, ell, engine=eng)
        steps = 20

        def step():
            cgef.compute_pi_(Pi, U, V, None, grid, ker, ell, workspace=ws, filter_plan=plan, rho0=1025.0)

        el, _ = timed_steps(step, steps, 3)
        r = {"engine": plan.engine, "ms_per_step": el * 1e3 / steps, "value": grid.npoints * steps / el,
             "unit": "point-scales/s"}
        if plan.engine in refs:
            ref = refs[plan.engine]
            got = Pi.cpu().numpy()
            r["rel_l2_vs_fp64_oracle"] = float(np.sqrt(np.sum((got - ref) ** 2) / np.sum(ref ** 2)))
        out[eng] = r
        log(f"cartesian 256^2 {plan.engine}: {r['ms_per_step']:.3f} ms/step")
    return out


def large_cartesian_leg(args):
    """HBM-bound form of the separable engine: 4096² doubly periodic Cartesian FP32, Gaussian α = 6,
    ℓ = 8Δx (33-tap row and column passes), the 256² generator's spectrum. The two filter kernels
    (row pass forming the five planes, column pass with the τ/ū epilogue) are timed together with
    HIP events; their compulsory HBM traffic per point is u, v (8 B) + mask (1 B) + five FP64 planes
    written and read (80 B) + ū (16 B) and τ (12 B) written = 117 B."""
    import numpy as np
    import torch

    import cgef
    from cgef import synthetic as S

    dev = "cuda:0"
    N, L = 4096, 1.6e7
    dx = L / N
    u, v, xs = S.cartesian_turbulence(N=N, L=L, seed=0, kmax=1300)
    grid = cgef.StructuredGrid(cgef.CartesianGeometry(), xs, xs, periodic=(True, True))
    ker = cgef.GaussianKernel(6.0)
    U = torch.tensor(u, dtype=torch.float32, device=dev)
    V = torch.tensor(v, dtype=torch.float32, device=dev)
    del u, v
    Pi = torch.empty_like(U)
    ws = cgef.PiWorkspace(grid, (), torch.float32)
    plan = cgef.plan_filter(grid, ker, 8 * dx)
    steps = 10

    def step():
        cgef.compute_pi_(Pi, U, V, None, grid, ker, 8 * dx, workspace=ws, filter_plan=plan)

    el, kms = timed_steps(step, steps, 2, plan)
    alg = 117.0 * grid.npoints
    out = {"workload": "cartesian 4096x4096 doubly periodic FP32, Gaussian alpha=6, l=8dx, separable",
           "engine": plan.engine, "ms_per_step": el * 1e3 / steps, "value": grid.npoints * steps / el,
           "unit": "point-scales/s", "filter_ms": kms, "filter_algorithmic_bytes": alg,
           "filter_hbm_gbs": alg / (kms * 1e-3) / 1e9, "filter_hbm_frac": alg / (kms * 1e-3) / 1e9 / HBM_PEAK_GBS}
    log(f"cartesian 4096^2 {plan.engine}: {out['ms_per_step']:.3f} ms/step, filter {kms:.3f} ms "
        f"({out['filter_hbm_gbs']:.0f} GB/s)")
    del U, V, Pi, ws, plan
    torch.cuda.empty_cache()
    return out


def spectral_leg(args):
    """method = Spectral() on doubly periodic Cartesian grids (the reference's FFTW extension):
    (a) Π on the 256² FP64 workload (Gaussian α = 6, ℓ = 8Δx, ρ₀ = 1025) with parity against the
    oracle's transform filter; (b) one FP64 top-hat filter at half-width 64 cells on 256², transform
    vs prefix sums (the reference's own comparison, Filtering.jl:3008-3011); (c) a 24-scale sweep
    (ℓ = 2Δx … 64Δx) sharing one analysis, vs the separable engine; (d) Π on 4096² FP32, with the
    column-transform kernel timed by HIP events."""
    import numpy as np
    import torch

    import cgef
    from cgef import synthetic as S

    dev = "cuda:0"
    out = {}
    N, L = 256, 1.0e6
    dx = L / N
    u, v, xs = S.cartesian_turbulence(N=N, L=L, seed=0)
    grid = cgef.StructuredGrid(cgef.CartesianGeometry(), xs, xs, periodic=(True, True))
    U, V = torch.tensor(u, device=dev), torch.tensor(v, device=dev)
    Pi = torch.empty_like(U)
    ws = cgef.PiWorkspace(grid, (), torch.float64)
    ker, ell = cgef.GaussianKernel(6.0), 8 * dx
    plan = cgef.plan_filter(grid, ker, ell, method=cgef.Spectral())
    steps = 20

    def step():
        cgef.compute_pi_(Pi, U, V, None, grid, ker, ell, workspace=ws, filter_plan=plan, rho0=1025.0)

    el, _ = timed_steps(step, steps, 3)
    r = {"workload": "cartesian 256x256 doubly periodic FP64, Gaussian alpha=6, l=8dx, rho0=1025, Spectral()",
         "engine": plan.engine, "ms_per_step": el * 1e3 / steps, "value": grid.npoints * steps / el,
         "unit": "point-scales/s"}
    if not args.no_cpu:
        from oracle import oracle as O
        og = O.Grid([xs, xs], periodic=[True, True], period=[L, L], uniform=[True, True])
        ref = 1025.0 * O.compute_pi(og, u, v, kind=O.GAUSSIAN, alpha=6.0, ell=ell, engine=O.ENGINE_SPECTRAL)[0]
        got = Pi.cpu().numpy()
        r["rel_l2_vs_fp64_oracle"] = float(np.sqrt(np.sum((got - ref) ** 2) / np.sum(ref ** 2)))
    out["pi_256"] = r
    log(f"spectral 256^2 Π: {r['ms_per_step']:.3f} ms/step")

    th = cgef.TopHatKernel()
    f = U.clone()
    g = torch.empty_like(f)
    r = {"workload": "one FP64 field, 256x256 doubly periodic, top-hat half-width 64 cells"}
    for eng, meth in (("spectral", cgef.Spectral()), ("prefixsum", None)):
        tp = cgef.plan_filter(grid, th, 128 * dx, method=meth)
        el, _ = timed_steps(lambda: cgef.filter_apply_(g, f, tp), 50, 5)
        r[f"{tp.engine}_ms"] = el * 1e3 / 50
    out["tophat_halfwidth64_256"] = r
    log(f"top-hat half-width 64: spectral {r['spectral_ms']:.3f} ms, prefix {r['prefixsum_ms']:.3f} ms")

    scales = list(np.geomspace(2 * dx, 64 * dx, 24))
    r = {"workload": "256x256 FP64 Gaussian alpha=6, 24 scales 2dx..64dx, coarse_grain"}
    for name, meth in (("spectral", cgef.Spectral()), ("separable", None)):
        plans = [cgef.plan_filter(grid, ker, s, method=meth) for s in scales]
        res = cgef.pipeline._allocate_result(grid, len(scales), torch.float64, dev)

        def sweep():
            cgef.coarse_grain_(res, U, V, None, grid, scales=scales, kernel=ker, filter_plans=plans,
                               spectrum=cgef.NoSpectrum())

        el, _ = timed_steps(sweep, 5, 1)
        r[f"{name}_ms_per_sweep"] = el * 1e3 / 5
        r[f"{name}_value"] = grid.npoints * len(scales) * 5 / el
    r["unit"] = "point-scales/s"
    out["sweep_24_scales_256"] = r
    log(f"24-scale sweep: spectral {r['spectral_ms_per_sweep']:.3f} ms, separable {r['separable_ms_per_sweep']:.3f} ms")
    del U, V, Pi, ws, plan, f, g

    N, L = 4096, 1.6e7
    dx = L / N
    u, v, xs = S.cartesian_turbulence(N=N, L=L, seed=0, kmax=1300)
    grid = cgef.StructuredGrid(cgef.CartesianGeometry(), xs, xs, periodic=(True, True))
    U = torch.tensor(u, dtype=torch.float32, device=dev)
    V = torch.tensor(v, dtype=torch.float32, device=dev)
    del u, v
    Pi = torch.empty_like(U)
    ws = cgef.PiWorkspace(grid, (), torch.float32)
    plan = cgef.plan_filter(grid, ker, 8 * dx, method=cgef.Spectral())
    steps = 10

    def step4k():
        cgef.compute_pi_(Pi, U, V, None, grid, ker, 8 * dx, workspace=ws, filter_plan=plan)

    el, kms = timed_steps(step4k, steps, 2, plan)
    # the column kernel reads and writes the five planes' half spectra once: 2 · 5 · (N/2+1) · N · 16 B
    alg = 2 * 5 * (N // 2 + 1) * N * 16
    out["pi_4096_fp32"] = {"workload": "cartesian 4096x4096 doubly periodic FP32, Gaussian alpha=6, l=8dx, Spectral()",
                           "engine": plan.engine, "ms_per_step": el * 1e3 / steps,
                           "value": grid.npoints * steps / el, "unit": "point-scales/s", "kernel": "k_spec_cols",
                           "kernel_ms": kms, "kernel_algorithmic_bytes": alg,
                           "kernel_hbm_gbs": alg / (kms * 1e-3) / 1e9}
    log(f"spectral 4096^2 Π: {el * 1e3 / steps:.3f} ms/step, column kernel {kms:.3f} ms")
    del U, V, Pi, ws, plan
    torch.cuda.empty_cache()
    return out


def closure_leg(args):
    """Learned τ closure on the doubly periodic k⁻³ box (1024², Gaussian α = 6, ℓ = 8Δx): exact (∇ū, τ)
    pairs from the filter for 64 training and 16 held-out seeds; the network learns what the gradient
    model σ²∇ū∇ūᵀ misses of τ, trained on the device for 2000 Adam steps at batch 16 (64² crops with
    their 8-cell context); then Π of the held-out fields against the exact Π, next to the gradient
    model alone, the exact pipeline's cost, and the exact-Π check with its 20 % fallback threshold."""
    import numpy as np
    import torch

    import cgef
    from cgef import closure as CL
    from cgef import synthetic as S

    N, ntrain, ntest = args.closure_n, 64, 16
    steps, batch, crop = args.closure_steps, 16, 64
    log(f"closure: generating {ntrain + ntest} exact pairs on the {N}x{N} box")
    t0 = time.perf_counter()
    feat, tau, grid, plan = CL.closure_dataset(list(range(1000, 1000 + ntrain + ntest)), N=N, ell_cells=8.0, device=0)
    torch.cuda.synchronize()
    t_data = time.perf_counter() - t0
    s2 = CL.gradient_model_sigma2(plan.kernel, plan.scale)
    ftr, ttr = feat[:ntrain], tau[:ntrain]
    fte, tte = feat[ntrain:].contiguous(), tau[ntrain:].contiguous()
    log(f"closure: data in {t_data:.1f} s; training {steps} steps, batch {batch}, {crop}^2 crops")
    t0 = time.perf_counter()
    model = CL.fit_closure(ftr, ttr, s2, steps=steps, batch=batch, crop=crop, log_every=max(1, steps // 8), seed=0,
                           log=log)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    net, hist = model.net, model.history
    h = crop + 2 * net.halo
    fwd_flops = batch * net.flops_per_sample(h, h)
    # forward and one training step's kernels on one batch, timed with events on the launch stream
    X = torch.empty((batch, 4, h, h), dtype=torch.bfloat16, device="cuda:0")
    T = torch.zeros((batch, 3, crop, crop), dtype=torch.float32, device="cuda:0")
    org = torch.zeros((batch, 3), dtype=torch.int32, device="cuda:0")
    CL.crop_(X, ftr, org, model.feat_scale)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    net.forward(X)
    torch.cuda.synchronize()
    reps = 20
    ev[0].record()
    for _ in range(reps):
        net.forward(X)
    ev[1].record()
    for _ in range(reps):
        net.forward(X)
        net.backward(T, accumulate_loss=False)
    ev[2].record()
    torch.cuda.synchronize()
    fwd_ms = ev[0].elapsed_time(ev[1]) / reps
    fb_ms = ev[1].elapsed_time(ev[2]) / reps
    # held-out evaluation
    t0 = time.perf_counter()
    pred = model.predict_pi(fte)
    torch.cuda.synchronize()
    t_inf = (time.perf_counter() - t0) / ntest
    exact = CL.pi_from_tau(fte, tte)
    corr, rel = CL.pi_errors(pred, exact)
    gcorr, grel = CL.pi_errors(CL.pi_from_tau(fte, CL.gradient_model_tau(fte, s2)), exact)
    # the exact pipeline's cost per field (compute_Π! on FP32 fields through the same plan), and the check
    u, v, _ = S.cartesian_turbulence(N=N, L=1.0e6, seed=1000 + ntrain)
    ut = torch.tensor(u, dtype=torch.float32, device="cuda:0")
    vt = torch.tensor(v, dtype=torch.float32, device="cuda:0")
    Pi = torch.empty_like(ut)

    def exact_step():
        cgef.compute_pi_(Pi, ut, vt, None, grid, plan.kernel, plan.scale, filter_plan=plan)

    t_exact = timed_steps(exact_step, 10, 2)[0] / 10
    checked = CL.LearnedClosure(model, grid, plan, threshold=0.2)
    out = torch.empty_like(ut)
    _, used_exact = checked.compute_pi_(out, ut, vt)
    log(f"closure: held-out Π corr {corr:.4f} rel-L2 {rel:.4f} (gradient model {gcorr:.4f} / {grel:.4f}); "
        f"checked error {checked.last_error:.4f}, exact used: {used_exact}")
    return {
        "workload": f"doubly periodic {N}x{N} k^-3 box, Gaussian alpha=6, l=8dx; net 4-64-64-64-3, 5x5, GELU, "
                    f"BF16 weights/activations, FP32 accumulation, learns tau minus the gradient model; Adam lr 1e-3, "
                    f"{steps} steps, batch {batch} {crop}^2 crops; {ntrain} training / {ntest} held-out seeds",
        "data_s": t_data, "train_s": t_train, "ms_per_train_step": 1e3 * t_train / steps,
        "final_train_mse": hist[-1][1] if hist else None,
        "forward_ms_per_batch": fwd_ms, "forward_tflops": fwd_flops / (fwd_ms * 1e-3) / 1e12,
        "train_step_kernels_ms": fb_ms, "train_step_tflops": 3 * fwd_flops / (fb_ms * 1e-3) / 1e12,
        "heldout_pi_corr": corr, "heldout_pi_rel_l2": rel,
        "gradient_model_pi_corr": gcorr, "gradient_model_pi_rel_l2": grel,
        "check_rel_l2": checked.last_error, "check_threshold": 0.2, "check_used_exact": used_exact,
        "inference_ms_per_field": 1e3 * t_inf, "exact_pi_ms_per_field": 1e3 * t_exact, "exact_engine": plan.engine,
    }


def volume_leg(args):
    """True 3-D: 256×256×64, Δx = Δy = 1, Δz_k = Δz₁·1.03^(k−1) with Δz = 4 at mid-depth (so the
    top-hat ball of radius ℓ/2 = 8 spans ±2 levels there), x/y periodic, z bounded, FP32, top-hat
    ℓ = 16Δx, ZeroFill, no mask. Parity against the oracle's 3-D ball engine."""
    import numpy as np
    import torch

    import cgef
    from cgef import synthetic as S

    dev = "cuda:0"
    u, v, w = S.volume_turbulence()
    Nz, Ny, Nx = u.shape
    xs, ys = np.arange(Nx, dtype=np.float64), np.arange(Ny, dtype=np.float64)
    zs = S.stretched_levels(Nz)
    grid = cgef.StructuredGrid(cgef.CartesianGeometry(), xs, ys, zs, periodic=(True, True, False))
    ker, ell = cgef.TopHatKernel(), 16.0
    U, V, W = (torch.tensor(a, dtype=torch.float32, device=dev) for a in (u, v, w))
    Pi = torch.empty_like(U)
    ws = cgef.PiWorkspace(grid, (), torch.float32)
    plan = cgef.plan_filter(grid, ker, ell)
    steps = 10

    def step():
        cgef.compute_pi_(Pi, U, V, W, grid, ker, ell, workspace=ws, filter_plan=plan)

    el, _ = timed_steps(step, steps, 2)
    out = {"workload": "3-D 256x256x64 stretched z (x1.03), FP32, top-hat l=16dx, x/y periodic",
           "engine": plan.engine, "ms_per_step": el * 1e3 / steps, "value": grid.npoints * steps / el,
           "unit": "point-scales/s", "mean_taps": plan.info.mean_taps}
    log(f"volume 256x256x64 {plan.engine}: {out['ms_per_step']:.3f} ms/step")
    if not args.no_cpu:
        from oracle import oracle as O
        og = O.Grid([xs, ys, zs], periodic=[True, True, False], period=[float(Nx), float(Ny), 0.0],
                    uniform=[True, True, False])
        ncores = min(len(os.sched_getaffinity(0)), 16)
        t0 = time.perf_counter()
        ref = O.compute_pi(og, *(a.cpu().double().numpy() for a in (U, V, W)), kind=O.TOPHAT, alpha=0.0, ell=ell,
                           engine=O.ENGINE_ND3, nthreads=ncores)[0]
        out["cpu_s"] = time.perf_counter() - t0
        out["cpu_cores"] = ncores
        got = Pi.cpu().double().numpy()
        out["rel_l2_vs_fp64_oracle"] = float(np.sqrt(np.sum((got - ref) ** 2) / np.sum(ref ** 2)))
        out["max_abs_err"] = float(np.max(np.abs(got - ref)))
    del U, V, W, Pi, ws, plan
    torch.cuda.empty_cache()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--engine", default="auto", help="auto (ring) | banded | baseline")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-extra", action="store_true", help="skip the secondary GPU engines")
    ap.add_argument("--no-batched", action="store_true", help="skip the 8-snapshot × 24-scale batched leg")
    ap.add_argument("--only", default=None,
                    help="comma-separated extra legs to run after the headline step (batched, tophat, cartesian, "
                         "cartesian4096, volume, spectral, closure); overrides --no-batched / --no-configs")
    ap.add_argument("--closure-steps", type=int, default=2000, help="Adam steps of the learned-closure leg")
    ap.add_argument("--closure-n", type=int, default=1024, help="box size of the learned-closure leg")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the top-hat sphere, Cartesian 256² and 4096², volume, spectral and learned-closure legs")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run one 512 MiB → 512 MiB FP32 elementwise copy first, so that hardware counter runs "
                         "can scale FETCH_SIZE / WRITE_SIZE by a kernel of known traffic")
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
    if args.copy_calibration:
        a = torch.ones(COPY_CAL_BYTES // 4, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        torch.mul(a, 1.0, out=b)
        torch.cuda.synchronize()
        del a, b
        log(f"copy calibration: one elementwise kernel reading and writing {COPY_CAL_BYTES} bytes each")
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
        nf = info.ring_pi_fields
        mh, noff, nrows = info.ring_modes, info.ring_offsets, grid.shape[1]
        traffic_alg = mh * noff * info.nbands * 8 + 2 * nf * mh * nrows * 16
        achieved = traffic_alg / (kern_ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic("k_contract"), "kernel": "k_contract",
                "kernel_ms": kern_ms, "algorithmic_bytes_per_launch": traffic_alg}
    else:
        flops = 2.0 * 9.0 * info.active_taps
        achieved = flops / (kern_ms * 1e-3) / 1e12
        roof = {"bound": "mfma", "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / FP32_PEAK_TFLOPS, "traffic": None, "kernel": "k_pi_filter", "kernel_ms": kern_ms,
                "algorithmic_flops_per_launch": flops}

    extra = {}
    # the direct sum's algorithmic flops (2 per in-support tap per filtered field, 9 fields on the
    # sphere) delivered per second by the measured engine
    extra["direct_sum_equivalent_tflops"] = 2.0 * 9.0 * info.active_taps / (ms_step * 1e-3) / 1e12
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

    legs = {"batched": ("batched_spectrum", lambda: batched_leg(args, grid, mask, lam, phi)),
            "tophat": ("tophat_sphere", lambda: tophat_leg(args, grid, lam, phi, mask, u32, v32)),
            "cartesian": ("cartesian_256", lambda: cartesian_leg(args)),
            "cartesian4096": ("cartesian_4096_separable", lambda: large_cartesian_leg(args)),
            "volume": ("volume_256x256x64", lambda: volume_leg(args)),
            "spectral": ("spectral", lambda: spectral_leg(args)),
            "closure": ("learned_closure", lambda: closure_leg(args))}
    if args.only is not None:
        chosen = [x for x in args.only.split(",") if x]
        unknown = [x for x in chosen if x not in legs]
        if unknown:
            print(f"bench.py: unknown leg(s) {unknown}; choose from {sorted(legs)}", file=sys.stderr)
            sys.exit(2)
    else:
        chosen = ([] if args.no_batched else ["batched"]) + \
                 ([] if args.no_configs else ["tophat", "cartesian", "cartesian4096", "volume", "spectral", "closure"])
    for name in chosen:
        key, run = legs[name]
        extra[key] = run()

    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        ncores = min(len(os.sched_getaffinity(0)), 16)
        log(f"FP64 CPU oracle on {ncores} threads (full field, one scale, repeated to >= 10 s)")
        og = O.Grid([lam, phi], spherical=True, R=6.371e6, periodic=[True, False], period=[2 * np.pi, 0.0],
                    mask=mask, uniform=[True, True])
        uin, vin = u32.cpu().double().numpy(), v32.cpu().double().numpy()
        fp = O.Footprint(og, O.GAUSSIAN, 4.0, SCALE)
        ncall, tc = 0, 0.0
        while tc < 10.0 and ncall < 20:
            t2 = time.perf_counter()
            ref, ub, vb, _ = O.compute_pi(og, uin, vin, kind=O.GAUSSIAN, alpha=4.0, ell=SCALE, fp=fp, nthreads=ncores)
            tc += time.perf_counter() - t2
            ncall += 1
        cpu = {"value": npts * ncall / tc, "unit": "point-scales/s", "cores": ncores, "kind": "port",
               "sample": f"full 1440x720 field at one scale, {ncall} calls in {tc:.1f} s"}
        diff = Pi_gpu - ref
        extra["rel_l2_vs_fp64_oracle"] = float(np.sqrt(np.sum(diff ** 2) / np.sum(ref ** 2)))
        extra["max_abs_err"] = float(np.max(np.abs(diff)))
        extra["rel_max_err"] = float(np.max(np.abs(diff)) / np.max(np.abs(ref)))
        # error table: coast band (footprint reaches land) vs open-ocean interior
        frac = O.apply_banded(og, fp, O.ZEROFILL, [mask.astype(np.float64)])[0]
        coast = (mask != 0) & (frac < 1.0 - 1e-9)
        interior = (mask != 0) & ~coast
        for name, sel in (("coast_band", coast), ("interior", interior)):
            extra[f"rel_l2_{name}"] = float(np.sqrt(np.sum(diff[sel] ** 2) / np.sum(ref[sel] ** 2)))
            extra[f"max_abs_err_{name}"] = float(np.max(np.abs(diff[sel])))
        extra["coast_band_points"] = int(coast.sum())
        Eo = O.energy(og, ub, vb)
        extra["energy_rel_err"] = float(abs(E_gpu - Eo) / abs(Eo))
        extra["speedup_vs_cpu"] = value / cpu["value"]
        log(f"parity: rel-L2 {extra['rel_l2_vs_fp64_oracle']:.3e}, CPU {tc / ncall:.2f} s per call")
        t2 = time.perf_counter()
        O.compute_pi(og, uin, vin, kind=O.GAUSSIAN, alpha=4.0, ell=SCALE, fp=fp, nthreads=1)
        t1 = time.perf_counter() - t2
        extra["cpu_baseline_1core"] = {"value": npts / t1, "unit": "point-scales/s", "cores": 1,
                                       "sample": f"full 1440x720 field at one scale, 1 call in {t1:.1f} s"}

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
        "dtype": "f32" if plan.engine == "banded" else "f64",
        "data": "synthetic",
        "config": {"workload": "spherical 0.25deg 1440x720 FP32, Gaussian alpha=4 haversine, l=200km, 30% land "
                               "ZeroFill, 1 scale, 1 snapshot", "engine": plan.engine,
                   "points": npts, "scales": 1, "snapshots": 1, "io_dtype": "f32", "parallelism": "1 GPU"},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    out.update(extra)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
