"""Benchmark: VI gradient evaluations per second of the X-ray forward model.

Workload (one GPU): sky 1024 x 1024 x 4 energies from the spectral GP
(xi_ref, xi_alpha on a 1024^2 grid, power-law slope -4, fluctuations 1,
offset -2), one module, linpatch_convolve with 8 x 8 patches and a 64-px
margin (FFT length 384), per-patch Gaussian+Moffat PSFs (beta 2.5, core
sigma 1-4 px), exposure ramp 5e2-2e3 s with 5 % of pixels masked, Poisson
counts.  One step = one gradient evaluation (GP forward, PSF forward,
exposure, Poisson NLL in FP64, PSF adjoint, GP adjoint) with inputs resident
in HBM.  The mixed path computes in FP32 with FP64 reductions; beside it run
the same fused kernels in FP64, the plain unfused FP64 GPU path (torch +
hipFFT + autograd, the reference's op order) and the float64 CPU oracle on
the host cores.  Prints one JSON line.
"""
import argparse
import json
import os
import statistics
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "VI grad evals/sec (resp+adjoint+NLL) 1024²×4E; mixed vs FP64 @1e-5 rel err"
HBM_PEAK = 8.0e12
FP32_PEAK = 157.3e12


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--H", type=int, default=1024)
    ap.add_argument("--E", type=int, default=4)
    ap.add_argument("--n", type=int, default=8)
    ap.add_argument("--margin", type=int, default=64)
    ap.add_argument("--cpu-seconds", type=float, default=25.0,
                    help="budget of the CPU-oracle sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--profile-only", action="store_true",
                    help="only run the mixed gradient steps (for rocprofv3)")
    ap.add_argument("--samples", type=int, default=8,
                    help="VI samples per call of the batched gradient leg (0: skip)")
    ap.add_argument("--tile", default="auto",
                    help="tile divisors of the column and row passes: the autotuner's pick "
                         "(auto) or fixed ones, 'C' or 'C,R' with C, R in {1, 2, 4} "
                         "(profiling runs repeat the pick without the tuner's own launches)")
    ap.add_argument("--sweep", action="store_true",
                    help="time every benchmark configuration instead of the 1024^2 x 4E line")
    ap.add_argument("--transform", default="fft", choices=("fft", "mfma_bf16x3"),
                    help="column-pass transform of the timed FP32 path (with --profile-only); "
                         "the full run times both at FFT length 384")
    ap.add_argument("--profile-samples", type=int, default=1,
                    help="with --profile-only: S > 1 runs the batched gradient of S VI-sample "
                         "positions per call instead of single evaluations")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="with --profile-only: first copy a buffer of this many MiB "
                         "(a known-size reference for the HBM byte counters)")
    return ap.parse_args()


def build_inputs(ja, H, E, n, m, seed=2):
    psf_windows = ja.synthetic.psf_windows
    rng = np.random.default_rng(seed)
    amp_ref = ja.powerlaw_amplitude(H)
    amp_alpha = 0.5 * amp_ref
    energies = np.array([1.0, 1.6, 2.5, 4.0][:E]) if E <= 4 else np.geomspace(1, 4, E)
    dnu = np.log(energies / energies[0])
    w = 2 * (H // n)
    win = psf_windows(rng, n, E, w, sigma=(1.0, 4.0), moffat=2.5)
    ramp = np.linspace(5e2, 2e3, H)[None, None, :]
    expo = np.broadcast_to(ramp, (E, H, H)).copy()
    expo[rng.random(expo.shape) < 0.05] = 0.0
    keep = expo > 0
    xi_true = rng.normal(size=(H, H))
    xa_true = rng.normal(size=(H, H))
    xi = xi_true + 0.1 * rng.normal(size=(H, H))
    xa = xa_true + 0.1 * rng.normal(size=(H, H))
    return dict(amp_ref=amp_ref, amp_alpha=amp_alpha, dnu=dnu, zm=-2.0, mu=-1.0, win=win,
                expo=expo, keep=keep, xi_true=(xi_true, xa_true), xi=(xi, xa), rng=rng,
                H=H, E=E, n=n, m=m, pixel_area=1.0)


def make_model(ja, torch, inp, dtype, counts, max_samples=1):
    H, E = inp["H"], inp["E"]
    sky = ja.CorrelatedFieldSky(H, (H, H), inp["amp_ref"], inp["amp_alpha"], zm=inp["zm"],
                                mu=inp["mu"], dnu=inp["dnu"], dtype=dtype)
    psf = ja.LinpatchOperator((E, H, H), inp["n"], inp["m"], dtype=dtype)
    psf.set_kernel(inp["win"], layout="window")
    return ja.XrayModel(sky, psf, inp["pixel_area"], inp["expo"][None], counts[None],
                        max_samples=max_samples)


def draw_counts(ja, torch, inp):
    """Poisson counts from lam at the true position (FP64 GPU forward)."""
    H, E = inp["H"], inp["E"]
    sky = ja.CorrelatedFieldSky(H, (H, H), inp["amp_ref"], inp["amp_alpha"], zm=inp["zm"],
                                mu=inp["mu"], dnu=inp["dnu"], dtype=torch.float64)
    s = sky(*[torch.as_tensor(v, device="cuda") for v in inp["xi_true"]])
    psf = ja.LinpatchOperator((E, H, H), inp["n"], inp["m"], dtype=torch.float64)
    psf.set_kernel(inp["win"], layout="window")
    lam = (psf.forward(s, scale=inp["pixel_area"]).cpu().numpy() * inp["expo"])
    counts = np.where(inp["keep"], inp["rng"].poisson(np.maximum(lam, 0)), -1).astype(np.int32)
    return counts


def time_steps(torch, fn, steps, warmup):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def window_rows_in_image(H, n, m):
    """Rows of the L-row windows (all n x n patches) that fall inside the
    image: the only ones the forward output and the adjoint input need."""
    dx = H // n
    h, w = dx // 2, 2 * dx
    L, lim = w + 2 * m, H + 2 * (dx // 2) - w
    tot = 0
    for ix in range(n):
        s = min(ix * dx, lim) - h - m
        tot += max(0, min(s + L, H) - max(s, 0))
    return tot * n


def kernel_bytes(H, E, n, m, M=1):
    """Bytes each kernel must move per launch given the three-pass design
    (inputs read once, outputs written once; FP32, complex = 8 B).  The
    forward output and the adjoint input of the column passes are the window
    rows inside the image (the others are never read, or zero)."""
    dx = H // n
    w, L = 2 * dx, 2 * dx + 2 * m
    L2 = L // 2  # stored half-spectrum columns (DC and Nyquist packed)
    J, P = M * E, n * n
    img = J * H * H * 4
    c = 8
    rows_in = window_rows_in_image(H, n, m)
    r1_fwd = J * P * w * L2 * c
    r2_fwd = J * rows_in * L2 * c
    r1_adj = J * rows_in * L2 * c
    r2_adj = J * P * w * L2 * c
    kt = M * P * E * L2 * L * c
    gp = H * H * c
    return {
        "gp_cols_fwd": 4 * H * H * 4 + gp,
        "gp_rows_fwd": gp + E * H * H * 4,
        "lp_rows_fwd": E * H * H * 4 + r1_fwd,
        "lp_cols_fwd": r1_fwd + kt + r2_fwd,
        "lp_fwd_out_poisson": r2_fwd + 4 * img,
        "lp_rows_adj": img + r1_adj,
        "lp_cols_adj": r1_adj + kt + r2_adj,
        "lp_adj_out_gp": r2_adj + E * H * H * 4 + gp,
        "gp_cols_adj": gp + gp,
        "gp_rows_adj": gp + 4 * H * H * 4,
    }


TRAFFIC_FILE = ROOT / "profiles" / "grad_eval_1024x4E_L384_hbm_traffic.txt"


def measured_traffic(kernel, H, E, n, m):
    """HBM bytes per dispatch of ``kernel`` from the committed rocprofv3 PMC
    summary (FETCH_SIZE and WRITE_SIZE, scaled by a known-size copy), for the
    default workload only; None otherwise.  ``__total__`` sums the kernels
    of one evaluation."""
    if (H, E, n, m) != (1024, 4, 8, 64) or not TRAFFIC_FILE.exists():
        return None
    rows = {}
    for line in TRAFFIC_FILE.read_text().splitlines():
        parts = line.split()
        if parts and not line.startswith("#") and len(parts) > 1:
            try:
                rows[parts[0]] = float(parts[-1])
            except ValueError:
                continue
    if kernel == "__total__":
        return sum(v for k, v in rows.items() if k in kernel_bytes(H, E, n, m)) or None
    return rows.get(kernel)


def eval_bytes(H, E, n, m, M=1):
    """B_alg of one gradient evaluation: 4 B x [2 |xi| + M E P w^2 + 2 M E H W]."""
    w = 2 * (H // n)
    xi = 2 * H * H
    return 4 * (2 * xi + M * E * n * n * w * w + 2 * M * E * H * H)


def fft_flops(H, E, n, m, M=1):
    """Nominal FLOPs: 2.5 N log2 N per real 2-D transform, 4 per patch-energy,
    6 L (L/2+1) per spectrum multiply (x2), GP 2 x 2.5 N log2 N per field x 2."""
    L = 2 * (H // n) + 2 * m
    pe = M * E * n * n
    t = 2.5 * L * L * np.log2(L * L)
    gp = 2 * 2 * 2.5 * H * H * np.log2(H * H)
    return pe * (4 * t + 2 * 6 * L * (L // 2 + 1)) + gp


def cpu_model():
    """The host CPU's model name (first /proc/cpuinfo entry)."""
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(inp, counts, budget):
    from oracle import convolve as ocv
    from oracle import model as om
    H, E = inp["H"], inp["E"]
    prob = om.Problem(shape=(E, H, H), distances=(1.0, 1.0, 1.0), n=inp["n"], margin=inp["m"],
                      pixel_area=inp["pixel_area"], amp_ref=inp["amp_ref"],
                      amp_alpha=inp["amp_alpha"], zm=inp["zm"], mu_alpha=inp["mu"],
                      dnu=inp["dnu"])
    keep = inp["keep"]
    prob.modules = [om.Module(window=inp["win"], exposure=inp["expo"], keep=keep,
                              data=counts[keep].astype(np.float64))]
    times = []
    t_start = time.perf_counter()
    result = None
    while True:
        t0 = time.perf_counter()
        result = om.nll_and_grad(prob, *inp["xi"])
        times.append(time.perf_counter() - t0)
        if time.perf_counter() - t_start > budget or len(times) >= 6:
            break
    timed = times[1:] if len(times) > 1 else times
    med = statistics.median(timed)
    return dict(value=1.0 / med, unit="grad evals/s", cores=ocv._WORKERS, kind="port",
                cpu_model=cpu_model(),
                sample=f"{len(timed)} full gradient evaluations of the same workload "
                       f"(float64 numpy/scipy oracle, scipy.fft workers={ocv._WORKERS}), "
                       f"median {med:.2f} s"), result


# Benchmark configurations: grid, energies, modules, patches, margin,
# PSF core sigma range (Gaussian, + Moffat beta), exposure range, mask
# fraction, samples per batched call.
CONFIGS = {
    "small_128x1E": dict(H=128, E=1, M=1, n=4, m=16, sigma=(1.0, 3.0), moffat=None,
                         expo=(1e3, 1e3), mask=0.10, S=0, seed=1),
    "adjoint_1024x1E": dict(H=1024, E=1, M=1, n=8, m=64, sigma=(1.0, 4.0), moffat=2.5,
                            expo=(5e2, 2e3), mask=0.05, S=0, seed=2),
    "modules_1024x3Ex7M": dict(H=1024, E=3, M=7, n=10, m=60, sigma=(2.0, 6.0), moffat=None,
                               expo=(1e2, 3e3), mask=0.08, S=0, seed=3),
    "grid_256x1E_S32": dict(H=256, E=1, M=1, n=6, m=22, sigma=(1.0, 3.0), moffat=None,
                            expo=(5e2, 2e3), mask=0.05, S=32, seed=4),
    "grid_512x1E_S32": dict(H=512, E=1, M=1, n=8, m=32, sigma=(1.0, 3.0), moffat=None,
                            expo=(5e2, 2e3), mask=0.05, S=32, seed=4),
    "grid_768x1E_S32": dict(H=768, E=1, M=1, n=12, m=48, sigma=(1.0, 3.0), moffat=None,
                            expo=(5e2, 2e3), mask=0.05, S=32, seed=4),
    "wide_2048x4E_S32": dict(H=2048, E=4, M=1, n=16, m=64, sigma=(1.0, 4.0), moffat=2.5,
                             expo=(5e2, 2e3), mask=0.05, S=32, seed=5),
    "grad_1024x4E_S8": dict(H=1024, E=4, M=1, n=8, m=64, sigma=(1.0, 4.0), moffat=2.5,
                            expo=(5e2, 2e3), mask=0.05, S=8, seed=2),
}


def sweep(args):
    """One JSON line per configuration: forward (sky + PSF + exposure),
    gradient evaluation, and for S > 0 the batched metric products and
    gradients, all with resident inputs; a one-evaluation CPU-oracle sample
    where it takes under ~10 s."""
    import torch
    import __graft_entry__ as ge
    ja = ge.load_package()
    powerlaw_amp, psf_windows = ja.powerlaw_amplitude, ja.synthetic.psf_windows
    uneven = []
    for name, c in CONFIGS.items():
        H, E, M, n, m = c["H"], c["E"], c["M"], c["n"], c["m"]
        rng = np.random.default_rng(c["seed"])
        spectral = E > 1
        amp_ref = powerlaw_amp(H)
        amp_alpha = 0.5 * amp_ref if spectral else None
        dnu = np.log(np.geomspace(1.0, 4.0, E)) if spectral else np.zeros(1)
        w = 2 * (H // n)
        win = np.stack([psf_windows(rng, n, E, w, sigma=c["sigma"], moffat=c["moffat"],
                                    rotation=k * np.pi / 7) for k in range(M)])
        ramp = np.linspace(c["expo"][0], c["expo"][1], H)[None, None, None, :]
        expo = np.broadcast_to(ramp, (M, E, H, H)).copy()
        expo[rng.random(expo.shape) < c["mask"]] = 0.0
        keep = expo > 0
        xt = rng.normal(size=(H, H))
        xat = rng.normal(size=(H, H)) if spectral else None

        def model_for(dtype, S=1, counts=None):
            sky = ja.CorrelatedFieldSky(H, (H, H), amp_ref, amp_alpha, zm=-2.0,
                                        mu=-1.0 if spectral else 0.0, dnu=dnu, dtype=dtype)
            psf = ja.LinpatchOperator((E, H, H), n, m, n_modules=M, dtype=dtype)
            psf.set_kernel(win if M > 1 else win[0], layout="window")
            cnt = counts if counts is not None else np.zeros(keep.shape, np.int32)
            return sky, psf, ja.XrayModel(sky, psf, 1.0, expo, cnt, max_samples=max(S, 1))

        # counts from the true position (FP64 forward)
        sky64, psf64, _ = model_for(torch.float64)
        s_true = sky64(torch.as_tensor(xt, device="cuda").to(sky64.dtype),
                       None if xat is None else torch.as_tensor(xat, device="cuda").to(sky64.dtype))
        lam = psf64.forward(s_true).reshape((M, E, H, H)) * torch.as_tensor(expo, device="cuda")
        counts = np.where(keep, ja.poisson_sample(lam, seed=c["seed"]).cpu().numpy(), -1)
        del sky64, psf64, s_true, lam
        xi = xt + 0.1 * rng.normal(size=(H, H))
        xa = xat + 0.1 * rng.normal(size=(H, H)) if spectral else None
        # the gradient of the FP64 build of the same kernels, for the error column
        _, _, m64 = model_for(torch.float64, 1, counts.astype(np.int32))
        n64, g64r, g64a = m64.nll_and_grad(torch.as_tensor(xi, device="cuda"),
                                           None if xa is None else torch.as_tensor(xa, device="cuda"))
        g64 = np.concatenate([g64r.cpu().numpy().ravel()] +
                             ([g64a.cpu().numpy().ravel()] if spectral else []))
        del m64, g64r, g64a
        torch.cuda.empty_cache()
        sky, psf, model = model_for(torch.float32, c["S"], counts.astype(np.int32))
        xr_t = torch.as_tensor(xi, dtype=torch.float32, device="cuda")
        xa_t = torch.as_tensor(xa, dtype=torch.float32, device="cuda") if spectral else None
        ex_t = torch.as_tensor(expo, dtype=torch.float32, device="cuda")
        steps = max(5, args.steps // (4 if H >= 2048 else 1))

        def fwd():
            return psf.forward(sky(xr_t, xa_t)).reshape(ex_t.shape) * ex_t

        out = model._grads()
        n32, g32r, g32a = model.nll_and_grad(xr_t, xa_t)
        g32 = np.concatenate([g32r.double().cpu().numpy().ravel()] +
                             ([g32a.double().cpu().numpy().ravel()] if spectral else []))
        rec = {"config": name, "H": H, "E": E, "M": M, "n": n, "margin": m,
               "L": 2 * (H // n) + 2 * m,
               "grad_rel_err_fp32_vs_fp64": rel(g32, g64), "grad_nmax_err_fp32_vs_fp64":
               nmax(g32, g64), "nll_rel_err_fp32_vs_fp64": abs(n32 - n64) / abs(n64),
               "forward_ms": 1e3 * time_steps(torch, fwd, steps, 2),
               "grad_eval_ms": 1e3 * time_steps(
                   torch, lambda: model.nll_and_grad_device(xr_t, xa_t, out=out), steps, 2)}
        # column tile picked by the autotuner for single evaluations
        tune = model.autotune(samples=1, reps=10)
        rec.update({"tile_ms": tune["ms"], "col_div": tune["col_div"], "row_div": tune["row_div"],
                    "grad_eval_ms_tuned": 1e3 * time_steps(
                        torch, lambda: model.nll_and_grad_device(xr_t, xa_t, out=out), steps, 2)})
        if c["S"]:
            S = c["S"]
            tune = model.autotune(samples=S, reps=3)
            rec.update({"tile_ms_batch": tune["ms"], "col_div_batch": tune["col_div"],
                        "row_div_batch": tune["row_div"]})
            model.nll_and_grad_device(xr_t, xa_t, out=out)
            v = torch.randn((S // 2, H, H), device="cuda", dtype=torch.float32)
            vr = torch.cat([v, -v])
            va = torch.cat([v, -v]) if spectral else None
            t_m = time_steps(torch, lambda: model.metric_batch(vr, va), max(3, steps // 8), 1)
            xb = xr_t[None] + 0.05 * torch.randn((S, H, H), device="cuda")
            xab = xa_t[None] + 0.05 * torch.randn((S, H, H), device="cuda") if spectral else None
            t_g = time_steps(torch, lambda: model.nll_and_grad_batch(xb, xab), max(3, steps // 8), 1)
            rec.update({"samples": S, "metric_batch_ms_per_product": 1e3 * t_m / S,
                        "grad_batch_ms_per_eval": 1e3 * t_g / S})
        gmf = None
        if psf.L == 384:
            # the split-BF16 matrix-core column transform on the same model
            psf.set_transform("mfma_bf16x3")
            nmf, gmr, gma = model.nll_and_grad(xr_t, xa_t)
            gmf = np.concatenate([gmr.double().cpu().numpy().ravel()] +
                                 ([gma.double().cpu().numpy().ravel()] if spectral else []))
            mrec = {"grad_rel_err_vs_fp64": rel(gmf, g64), "grad_nmax_err_vs_fp64": nmax(gmf, g64),
                    "nll_rel_err_vs_fp64": abs(nmf - n64) / abs(n64),
                    "grad_eval_ms": 1e3 * time_steps(
                        torch, lambda: model.nll_and_grad_device(xr_t, xa_t, out=out), steps, 2)}
            if c["S"]:
                S = c["S"]
                xb = xr_t[None] + 0.05 * torch.randn((S, H, H), device="cuda")
                xab = (xa_t[None] + 0.05 * torch.randn((S, H, H), device="cuda")
                       if spectral else None)
                mrec["grad_batch_ms_per_eval"] = 1e3 * time_steps(
                    torch, lambda: model.nll_and_grad_batch(xb, xab), max(3, steps // 8), 1) / S
            psf.set_transform("fft")
            rec["mfma_bf16x3"] = mrec
        if not args.no_cpu:
            # one evaluation of the float64 CPU oracle: timing and the error
            # of both GPU builds against it
            from oracle import model as om
            prob = om.Problem(shape=(E, H, H), distances=(1.0, 1.0, 1.0), n=n, margin=m,
                              pixel_area=1.0, amp_ref=amp_ref, amp_alpha=amp_alpha, zm=-2.0,
                              mu_alpha=-1.0 if spectral else 0.0, dnu=dnu)
            prob.modules = [om.Module(window=win[k], exposure=expo[k], keep=keep[k],
                                      data=counts[k][keep[k]].astype(np.float64))
                            for k in range(M)]
            t0 = time.perf_counter()
            no, (gor, goa) = om.nll_and_grad(prob, xi, xa)
            rec["cpu_oracle_grad_eval_ms"] = 1e3 * (time.perf_counter() - t0)
            go = np.concatenate([gor.ravel()] + ([goa.ravel()] if spectral else []))
            rec.update({"grad_rel_err_fp32_vs_oracle": rel(g32, go),
                        "grad_nmax_err_fp32_vs_oracle": nmax(g32, go),
                        "grad_rel_err_fp64_vs_oracle": rel(g64, go),
                        "nll_rel_err_fp32_vs_oracle": abs(n32 - no) / abs(no)})
            if gmf is not None:
                rec["mfma_bf16x3"].update({"grad_rel_err_vs_oracle": rel(gmf, go),
                                           "grad_nmax_err_vs_oracle": nmax(gmf, go)})
        print(json.dumps(rec), flush=True)
        if name.startswith("grid_"):
            # kept for the uneven-grid graph below, with its single-evaluation tile
            model.set_tile(rec["col_div"], rec["row_div"])
            uneven.append((name, model, xr_t, xa_t))
        del model, sky, psf
        torch.cuda.empty_cache()
    if len(uneven) > 1:
        print(json.dumps(uneven_graph(ja, torch, uneven, args.steps)), flush=True)
    del uneven
    print(json.dumps(dft_compare(ja, torch, args)), flush=True)


def uneven_graph(ja, torch, uneven, steps):
    """One gradient evaluation of each observation of an uneven-grid set:
    eager calls one after another, one captured graph per observation
    launched one after another, and one graph holding all of them as
    parallel branches (jubik_graph_capture_batch)."""
    outs = [mdl._grads() for _, mdl, _, _ in uneven]

    def eager():
        for (_, mdl, xr, xa), o in zip(uneven, outs):
            mdl.nll_and_grad_device(xr, xa, out=o)

    singles = [mdl.capture_grad(0) for _, mdl, _, _ in uneven]
    batch = ja.capture_grad_batch([(mdl, 0) for _, mdl, _, _ in uneven])
    for gs, (_, _, xr, xa) in zip(singles, uneven):
        gs.xi_ref.copy_(xr)
        if xa is not None:
            gs.xi_alpha.copy_(xa)
    for sl, (_, _, xr, xa) in zip(batch.slots, uneven):
        sl.xi_ref.copy_(xr)
        if xa is not None:
            sl.xi_alpha.copy_(xa)

    def graphs():
        for gs in singles:
            gs.launch()

    t_e = time_steps(torch, eager, steps, 2)
    t_s = time_steps(torch, graphs, steps, 2)
    t_b = time_steps(torch, batch.launch, steps, 2)
    return {"uneven_grids": [nm for nm, _, _, _ in uneven], "eager_ms": 1e3 * t_e,
            "graph_per_grid_ms": 1e3 * t_s, "one_graph_ms": 1e3 * t_b}


def dft_compare(ja, torch, args):
    """The 128-point DFT comparison of the 2048^2 x 4 configuration: 2^20
    vectors (1 GiB in, 1 GiB out) through the FP32 FFT and the split-BF16
    matrix-core DFT; accuracy against numpy in float64 on 4096 of them."""
    B = 1 << 20
    g = torch.Generator(device="cuda").manual_seed(0)
    x = torch.complex(torch.randn((B, 128), generator=g, device="cuda"),
                      torch.randn((B, 128), generator=g, device="cuda"))
    d = ja.DFT128()
    y = torch.empty_like(x)
    want = np.fft.fft(x[:4096].cpu().numpy().astype(np.complex128), axis=1)
    rec = {"dft128": True, "vectors": B}
    for method in ("fft", "mfma"):
        t = time_steps(torch, lambda: d(x, method, out=y), max(5, args.steps // 5), 2)
        got = y[:4096].cpu().numpy()
        rec[method] = {"ms": 1e3 * t, "vectors_per_s": B / t,
                       "hbm_GBps": 2 * x.numel() * 8 / t / 1e9,
                       "rel_err_vs_fp64": float(np.linalg.norm(got - want) / np.linalg.norm(want))}
    # 4 row tiles x 8 k-steps x 12 products of 32x32x16 per 32 vectors
    rec["mfma"]["tflops_matrix_core"] = B / 32 * 4 * 8 * 12 * (32 * 32 * 16 * 2) / (
        rec["mfma"]["ms"] * 1e-3) / 1e12
    return rec


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def nmax(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the hot path is sharded by module/energy/sample across GPUs by design, "
              "but only the one-GPU figure is measured)", file=sys.stderr)
        sys.exit(2)
    if args.sweep:
        sweep(args)
        return
    import torch
    import __graft_entry__ as ge
    ja = ge.load_package()
    from jubik_amd import _lib

    H, E, n, m = args.H, args.E, args.n, args.margin
    inp = build_inputs(ja, H, E, n, m)
    counts = draw_counts(ja, torch, inp)
    mixed = make_model(ja, torch, inp, torch.float32, counts)
    # tiles of the column and row passes picked by the autotuner (set-up)
    fixed = None if args.tile == "auto" else tuple(int(v) for v in args.tile.split(","))
    if fixed is None:
        tune = mixed.autotune(samples=1, reps=10)
    else:
        mixed.set_tile(*fixed)
        tune = {"col_div": fixed[0], "row_div": fixed[1] if len(fixed) > 1 else 1, "ms": None}
    xi32 = [torch.as_tensor(v, dtype=torch.float32, device="cuda") for v in inp["xi"]]
    out32 = mixed._grads()
    step = lambda: mixed.nll_and_grad_device(*xi32, out=out32)
    if args.profile_only:
        mixed.psf.set_transform(args.transform)
        if args.calib_copy_mib:
            src = torch.ones(args.calib_copy_mib << 18, dtype=torch.float32, device="cuda")
            dst = torch.empty_like(src)
            dst.copy_(src)
            torch.cuda.synchronize()
            del src, dst
        if args.profile_samples > 1:
            S = args.profile_samples
            bm = make_model(ja, torch, inp, torch.float32, counts, max_samples=S)
            bm.set_tile(*(fixed if fixed is not None else (1, 1)))
            brng = np.random.default_rng(17)
            xb = [torch.as_tensor(v[None] + 0.05 * brng.normal(size=(S,) + v.shape),
                                  dtype=torch.float32, device="cuda") for v in inp["xi"]]
            t = time_steps(torch, lambda: bm.nll_and_grad_batch(*xb), args.steps, args.warmup)
            print(json.dumps({"profile_only": True, "samples": S, "ms_per_call": t * 1e3}))
            return
        t = time_steps(torch, step, args.steps, args.warmup)
        print(json.dumps({"profile_only": True, "ms_per_step": t * 1e3}))
        return

    t_mixed = time_steps(torch, step, args.steps, args.warmup)

    # the same evaluation replayed from a captured HIP graph
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        step()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        step()
    t_graph = time_steps(torch, graph.replay, args.steps, args.warmup)
    del graph

    # host buffers (numpy positions in, numpy gradients out, the NLL read
    # back): the PCIe-inclusive rate of the drop-in functions, not the value
    xi_host = [None if v is None else np.ascontiguousarray(v, dtype=np.float32)
               for v in inp["xi"]]
    t_host = time_steps(torch, lambda: mixed.nll_and_grad(*xi_host), max(3, args.steps // 5), 2)

    # S VI-sample positions per call (jubik_model_grad_batch)
    batched = None
    if args.samples > 1:
        S = args.samples
        bm = make_model(ja, torch, inp, torch.float32, counts, max_samples=S)
        if fixed is None:
            tune_b = bm.autotune(samples=S, reps=3)
        else:
            bm.set_tile(*fixed)
            tune_b = {"col_div": fixed[0], "row_div": fixed[1] if len(fixed) > 1 else 1}
        brng = np.random.default_rng(17)
        xb = [torch.as_tensor(v[None] + 0.05 * brng.normal(size=(S,) + v.shape),
                              dtype=torch.float32, device="cuda") for v in inp["xi"]]
        t_b = time_steps(torch, lambda: bm.nll_and_grad_batch(*xb), max(3, args.steps // S), 2)
        batched = {"samples": S, "ms_per_call": t_b * 1e3, "evals_per_s": S / t_b,
                   "tile_divisors": [tune_b["col_div"], tune_b["row_div"]]}
        del bm, xb
        torch.cuda.empty_cache()

    nll32, g32r, g32a = mixed.nll_and_grad_device(*xi32)
    torch.cuda.synchronize()
    g32 = np.concatenate([g32r.double().cpu().numpy().ravel(), g32a.double().cpu().numpy().ravel()])
    nll32 = float(nll32.item())

    # the same evaluation with the split-BF16 matrix-core column transform
    mfma = None
    if mixed.psf.L == 384:
        mixed.psf.set_transform("mfma_bf16x3")
        t_mf = time_steps(torch, step, args.steps, args.warmup)
        nmf, gmr, gma = mixed.nll_and_grad_device(*xi32)
        torch.cuda.synchronize()
        gmf = np.concatenate([gmr.double().cpu().numpy().ravel(),
                              gma.double().cpu().numpy().ravel()])
        _lib.timing_enable(True)
        for _ in range(max(10, min(args.steps, 30))):
            step()
        torch.cuda.synchronize()
        rep_mf = _lib.timing_report()
        _lib.timing_enable(False)
        mixed.psf.set_transform("fft")
        mfma = {"ms_per_step": t_mf * 1e3, "evals_per_s": 1.0 / t_mf,
                "nll": float(nmf.item()), "grad": gmf,
                "column_pass_us_bracketed": {
                    k: 1e3 * tot / c for k, (c, tot, lo, hi) in rep_mf.items()
                    if k.startswith("lp_cols")}}

    # Per-kernel HIP-event timing on the launch stream (outside the timed
    # region).  With every launch bracketed the per-kernel times carry the
    # events' own cost, so they only rank the kernels; the dominant kernel is
    # then timed alone (events around its launches only, the rest of each
    # evaluation runs unbracketed).
    nprof = max(10, min(args.steps, 30))
    _lib.timing_enable(True)
    for _ in range(nprof):
        step()
    torch.cuda.synchronize()
    rep = _lib.timing_report()
    kernels = {k: {"launches": c, "avg_us_bracketed": 1e3 * tot / c}
               for k, (c, tot, lo, hi) in rep.items()}
    dom = max(rep.items(), key=lambda kv: kv[1][1])[0]
    _lib.timing_select(dom)
    for _ in range(nprof):
        step()
    torch.cuda.synchronize()
    rep_dom = _lib.timing_report()
    _lib.timing_select(None)
    _lib.timing_enable(False)
    dom_avg = rep_dom[dom][1] / rep_dom[dom][0] * 1e-3
    kb = kernel_bytes(H, E, n, m)
    # B_alg of one gradient evaluation; one launch of the dominant kernel
    # processes all J * P patch-energies of one evaluation.
    b_alg = eval_bytes(H, E, n, m)
    achieved = b_alg / dom_avg
    traffic_dom = measured_traffic(dom, H, E, n, m)
    traffic_eval = measured_traffic("__total__", H, E, n, m)

    # fused kernels in FP64
    f64 = make_model(ja, torch, inp, torch.float64, counts)
    xi64 = [torch.as_tensor(v, dtype=torch.float64, device="cuda") for v in inp["xi"]]
    o64 = f64._grads()
    t_f64 = time_steps(torch, lambda: f64.nll_and_grad_device(*xi64, out=o64),
                       max(5, args.steps // 5), 2)
    nll64, g64r, g64a = f64.nll_and_grad_device(*xi64)
    torch.cuda.synchronize()
    g64 = np.concatenate([g64r.cpu().numpy().ravel(), g64a.cpu().numpy().ravel()])
    nll64 = float(nll64.item())
    del f64

    plain = None
    if not args.no_plain:
        from plain_fp64 import PlainFP64Model
        w = 2 * (H // n)
        kfull = torch.zeros((n * n, E, H, H), dtype=torch.float64, device="cuda")
        c = (H - w) // 2
        kfull[..., c:c + w, c:c + w] = torch.as_tensor(inp["win"], device="cuda")
        kfull = torch.fft.ifftshift(kfull, dim=(-2, -1)).contiguous()
        res = {}
        for prep in (True, False):
            pm = PlainFP64Model(inp["amp_ref"], inp["amp_alpha"], inp["zm"], inp["mu"],
                                inp["dnu"], kfull, n, m, inp["pixel_area"], inp["expo"],
                                inp["keep"], counts[inp["keep"]].astype(np.float64),
                                prepare_each_call=prep)
            t = time_steps(torch, lambda: pm.nll_and_grad(*xi64), 3, 1)
            val, gr, ga = pm.nll_and_grad(*xi64)
            gp_ = np.concatenate([gr.cpu().numpy().ravel(), ga.cpu().numpy().ravel()])
            res["per_call_kernel_prep" if prep else "prepared_kernel"] = dict(
                ms_per_step=t * 1e3, evals_per_s=1.0 / t,
                grad_rel_err_vs_fused_fp64=rel(gp_, g64))
            del pm
        del kfull
        torch.cuda.empty_cache()
        plain = res

    cpu = None
    grad_vs_oracle = None
    if not args.no_cpu:
        cpu, (nll_o, (gor, goa)) = cpu_baseline(inp, counts, args.cpu_seconds)
        go = np.concatenate([gor.ravel(), goa.ravel()])
        grad_vs_oracle = {"mixed_rel": rel(g32, go), "mixed_nmax": nmax(g32, go),
                          "fused_fp64_rel": rel(g64, go), "nll_mixed_rel": abs(nll32 - nll_o) /
                          abs(nll_o), "nll_fp64_rel": abs(nll64 - nll_o) / abs(nll_o)}

    # practical HBM ceiling: a 1 GiB device-to-device copy (read + write)
    src = torch.empty(1 << 28, dtype=torch.float32, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    t_copy = time_steps(torch, lambda: dst.copy_(src), 20, 3)
    copy_bw = 2 * src.numel() * 4 / t_copy
    del src, dst

    value = 1.0 / t_mixed
    best_plain = max(plain.values(), key=lambda r: r["evals_per_s"]) if plain else None
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "grad evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_mixed * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{H}x{H}x{E}E spectral GP sky, 1 module, linpatch n={n} "
                               f"margin={m} (L={2 * (H // n) + 2 * m}), Gaussian+Moffat PSFs, "
                               "exposure 5e2-2e3 s, 5% mask, Poisson NLL (FP64 reduction)",
                   "precision": "FP32 FFTs/PSF products/exp, FP64 NLL reductions"},
        "autotune": {"column_tile_divisor": tune["col_div"], "row_tile_divisor": tune["row_div"],
                     "ms_per_forward_adjoint": tune["ms"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": traffic_dom, "kernel": dom,
                     "kernel_avg_us": dom_avg * 1e6, "alg_bytes_per_launch": b_alg,
                     "units_per_launch": f"one gradient evaluation ({E * n * n} patch-energies)",
                     "traffic_over_alg": (traffic_dom / b_alg) if traffic_dom else None},
        "eval_roofline": {"B_alg": b_alg, "achieved": b_alg / t_mixed,
                          "frac": b_alg / t_mixed / HBM_PEAK,
                          "traffic": traffic_eval,
                          "traffic_over_alg": (traffic_eval / b_alg) if traffic_eval else None,
                          "fft_flops": fft_flops(H, E, n, m),
                          "frac_fp32": fft_flops(H, E, n, m) / t_mixed / FP32_PEAK,
                          "kernel_sum_us_bracketed": sum(k["avg_us_bracketed"]
                                                         for k in kernels.values())},
        "kernel_streaming": {"kernel": dom, "design_bytes_per_launch": kb.get(dom, 0),
                             "GBps": kb.get(dom, 0) / dom_avg / 1e9,
                             "frac_of_peak": kb.get(dom, 0) / dom_avg / HBM_PEAK,
                             "device_copy_GBps": copy_bw / 1e9,
                             "frac_of_copy": kb.get(dom, 0) / dom_avg / copy_bw},
        "cpu_baseline": cpu,
        "fused_fp64": {"ms_per_step": t_f64 * 1e3, "evals_per_s": 1.0 / t_f64,
                       "grad_rel_err_mixed_vs_fp64": rel(g32, g64),
                       "grad_nmax_err_mixed_vs_fp64": nmax(g32, g64)},
        "plain_fp64": plain,
        "speedup_vs_plain_fp64": (value / best_plain["evals_per_s"]) if best_plain else None,
        "grad_err_vs_cpu_oracle": grad_vs_oracle,
        "hip_graph": {"ms_per_step": t_graph * 1e3, "evals_per_s": 1.0 / t_graph},
        "host_buffers": {"ms_per_step": t_host * 1e3, "evals_per_s": 1.0 / t_host,
                         "what": "numpy float32 positions in, numpy gradients out, NLL read "
                                 "back each call (PCIe both ways); not the bench value"},
        "batched_samples": batched,
        "kernels": kernels,
    }
    if mfma is not None:
        gmf = mfma.pop("grad")
        mfma.update({"grad_rel_err_vs_fp64": rel(gmf, g64), "grad_nmax_err_vs_fp64": nmax(gmf, g64),
                     "nll_rel_err_vs_fp64": abs(mfma["nll"] - nll64) / abs(nll64)})
        if not args.no_cpu:
            mfma.update({"grad_rel_err_vs_oracle": rel(gmf, go),
                         "grad_nmax_err_vs_oracle": nmax(gmf, go)})
        line["column_transform_mfma_bf16x3"] = mfma
    print(json.dumps(line))


if __name__ == "__main__":
    main()
