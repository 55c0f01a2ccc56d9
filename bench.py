"""Benchmark of the HIP rasterizer's hot path: one fwd+bwd of a 2048x2048 field of 200k Gaussians.

A step is one pass of the renderer's hot path over the resident field: tile binning (no host
synchronization), tiled forward, and the tiled backward of the default autograd path, against a
fixed synthetic dL/dout.  By default the backward recomputes the per-pixel normalizer with a
second sweep of the tile lists (a denominator-only sweep) instead of storing it, and reads the
forward's image; the variant that stores the normalizer (the reference's saved den/out) is timed
beside it (`--normalizer stored` makes it the step).  Prints one JSON line (metric, value, roofline of the dominant kernel, CPU
baseline).  `value` is the algorithmic HBM throughput of the whole step, by the workload's fixed
byte model:

    B_step = B_bin + B_fwd + B_bwd
    B_bin  = 52 N + 4 P + 4 (T+1)                     (means, conics, radii in; lists, offsets out)
    B_fwd  = 4 (T+1) + P (4 + R) + 12 H W             (one list sweep; RGB out)
    B_bwd  = 12 H W + 2 P (4 + R) + 36 P + 36 N       (grad in; two list sweeps; one 9-float
                                                        atomic set per pair; gradients)
with R = 44 bytes of Gaussian parameters per pair visit (mean 8, conic 12, color 12, opacity 4,
radii 8), N Gaussians, P (tile, Gaussian) pairs and T tiles.  The model is the same for either
normalizer mode: it is the workload's unit of work, so a faster mode shows as a shorter step.

P in `value` is counted at the reference's 16-pixel tiling whatever tile the renderer runs: the
workload's unit of work stays fixed, so a tiling with fewer pairs shows up as a shorter step, not
as fewer bytes.  The render tile (16 or 32 pixels) is the autotuner's choice on this field
(`--tile-size` overrides it); every candidate's fwd+bwd time is reported.

The roofline is reported for the step's dominant kernel, the backward (k_backward, or
k_backward_persistent under the persistent schedule), at the chosen tile: its own algorithmic
bytes per launch, with that tile's P,

    stored normalizer:     B_k = 28 H W + P (4 + R) + 36 P   (grad, den, image in; one sweep; slots)
    recomputed normalizer: B_k = 24 H W + 2 P (4 + R) + 36 P (grad, image in; two sweeps; slots)

(the 36 N gradient bytes belong to k_grad_reduce), over its average duration, measured by HIP
events that the library records on the launch stream around that kernel alone.  `traffic` is the kernel's
measured HBM bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, calibrated against a plain copy;
`--copy-probe` is that copy) when profiles/hbm_traffic.json holds a measurement of this workload
and kernel, else null.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload metric|single|stress]
                       [--tile-size auto|16|32|272|288]
       python bench.py --copy-probe MIB     (one device copy of MIB MiB, for counter calibration)
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import os
import platform
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

RECORD_BYTES = 44  # algorithmic bytes of one Gaussian's parameters
HBM_PEAK_GBS = 8000.0
# Vector FP32 issue rate of MI355X without packed math: 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz.
VALU_PEAK_LANE_OPS = 39.3e12
# Useful FP32 lane-operations per in-rectangle (pixel, Gaussian) visit, counted from the kernels'
# arithmetic: forward dx, dy, the quadratic (5), exp2, the denominator and 3 color FMAs = 12;
# backward's gradient visit 23 (the quadratic and exp2 again, dw 3, 4 color/opacity sums, 7 for
# the moment sums) plus the recomputed denominator's 8.
VISIT_OPS = {"fwd": 12, "bwd_stored": 23, "bwd_recompute": 31}
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic.json")

WORKLOADS = {
    "metric": ("fwd+bwd 2048x2048 RGB fp32, 200,000 Gaussians: scales log-U[0.5,12] px, aspect "
               "log-U[1,10] (major s, minor s/a), smoothed random angle field, opacity U[0.1,1], "
               "normalized, sigma_cutoff 3, seed 0", 2048, 2048, 200_000),
    "single": ("fwd+bwd 1024x1024 RGB fp32, 50,000 Gaussians, same recipe, seed 0",
               1024, 1024, 50_000),
    "stress": ("fwd+bwd 4096x4096 RGB fp32, 1,000,000 Gaussians: scales log-normal(ln 2, 1) "
               "clipped at 64 px, aspect log-U[1,20], seed 0", 4096, 4096, 1_000_000),
}


def rect_visits(means, radii, H: int, W: int) -> int:
    """E: the (pixel, Gaussian) visits of the whole image, the pixels of every Gaussian's clipped
    support rectangle (centre rounded half to even, +- radius, clipped; render.py:62-78)."""
    import torch

    m = means.double()
    r = radii.long()
    ok = torch.isfinite(m).all(dim=1)
    ix = torch.round(torch.nan_to_num(m[:, 0])).long()
    iy = torch.round(torch.nan_to_num(m[:, 1])).long()
    x0 = (ix - r[:, 0]).clamp(min=0)
    x1 = (ix + r[:, 0]).clamp(max=W - 1)
    y0 = (iy - r[:, 1]).clamp(min=0)
    y1 = (iy + r[:, 1]).clamp(max=H - 1)
    area = (x1 - x0 + 1).clamp(min=0) * (y1 - y0 + 1).clamp(min=0)
    return int((area * ok).sum().item())


def backward_kernel_bytes(P: int, HW: int, stored: bool) -> int:
    """Algorithmic bytes of one backward kernel launch (gradient slots out, not the reduction)."""
    R = RECORD_BYTES
    if stored:
        return 28 * HW + P * (4 + R) + 36 * P
    return 24 * HW + 2 * P * (4 + R) + 36 * P


def algorithmic_bytes(N: int, P: int, T: int, HW: int) -> dict[str, float]:
    R = RECORD_BYTES
    b_bin = 52 * N + 4 * P + 4 * (T + 1)
    b_fwd = 4 * (T + 1) + P * (4 + R) + 12 * HW
    b_bwd = 12 * HW + 2 * P * (4 + R) + 36 * P + 36 * N
    return {"bin": b_bin, "fwd": b_fwd, "bwd": b_bwd, "step": b_bin + b_fwd + b_bwd}


def make_field(workload: str, device):
    import torch

    from structsplat_amd import synth
    from structsplat_amd.gaussians import GaussianField

    if workload == "metric":
        f = synth.metric_field(0)
    elif workload == "cpu_reference":
        f = synth.cpu_reference_field(0)
    elif workload == "single":
        f = synth.single_field(0)
    else:
        f = synth.stress_field(0)
    logits = torch.logit(torch.tensor(f["opacity"], dtype=torch.float64)).float().numpy()
    field = GaussianField.from_numpy(f["means"], f["scales"], f["angles"], f["colors"],
                                     opacities=logits, device=device)
    with torch.no_grad():
        means = field.means.contiguous()
        conics = field.conics().contiguous()
        colors = field.colors.contiguous()
        radii = field.radii(3.0).contiguous()
        op = field.opacity_values().contiguous()
    return means, conics, colors, radii, op


def _median_seconds(fn, warmup: int, reps: int, sync=None) -> float:
    for _ in range(warmup):
        fn()
    if sync:
        sync()
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        if sync:
            sync()
        times.append(time.perf_counter() - t0)
    return statistics.median(times)


def _render_timings(fields, H: int, W: int, device, warmup: int, reps: int) -> dict:
    """Median seconds of the normalized render alone and of render + backward (dL/dout = 1/(3HW),
    the mean's gradient) of one field: on the CPU through the oracle's torch restatement of
    render.py (float32, render_chunk 512, torch autograd), on the GPU through render_hip_exact."""
    import torch

    from oracle import reference_render as R
    from structsplat_amd import render as HR

    means, conics, colors, radii, op = (t.to(device) for t in fields)
    grad = torch.full((H, W, 3), 1.0 / (3 * H * W), device=device)
    on_cpu = device == "cpu"
    sync = None if on_cpu else torch.cuda.synchronize

    def fwd():
        with torch.no_grad():
            if on_cpu:
                R.render(means, conics, colors, radii, H, W, chunk=512, opacities=op)
            else:
                HR.render_hip_exact(means, conics, colors, radii, H, W, opacities=op)

    def fwd_bwd():
        leaves = [t.detach().clone().requires_grad_() for t in (means, conics, colors, op)]
        if on_cpu:
            img = R.render(leaves[0], leaves[1], leaves[2], radii, H, W, chunk=512,
                           opacities=leaves[3])
        else:
            img = HR.render_hip_exact(leaves[0], leaves[1], leaves[2], radii, H, W,
                                      opacities=leaves[3])
        img.backward(grad)

    return {"fwd_s": _median_seconds(fwd, warmup, reps, sync),
            "fwd_bwd_s": _median_seconds(fwd_bwd, warmup, reps, sync)}


def cpu_baseline(workload: str, H: int, W: int, bytes_step: float) -> dict:
    """The reference CPU path timed on this host's cores beside the GPU, on identical seeded
    inputs: the oracle's torch restatement of render.py (_accumulate with render_chunk 512, torch
    autograd for the backward).  Two workloads, forward and forward + backward each:

    * the reference's CPU case, 256x256 / 2,000 Gaussians, seed 0: 3 warm-ups, median of 10;
    * the metric workload (2048x2048 / 200,000): 3 warm-ups, median of 3.

    `value` is the metric workload's fwd+bwd in the metric's unit (algorithmic GB/s); every
    timing sits next to the GPU's time for the same call (render_hip_exact with autograd) and
    the GPU/CPU speedup."""
    import torch

    threads = os.cpu_count() or 1
    torch.set_num_threads(threads)
    gpu = torch.device("cuda:0")
    out = {}
    for name, (h, w), warmup, reps in (("cpu_reference_256x256_2k", (256, 256), 3, 10),
                                       (f"{workload}_{W}x{H}", (H, W), 3, 3)):
        fields = tuple(t.cpu() for t in make_field("cpu_reference" if h == 256 else workload,
                                                   "cpu"))
        cpu = _render_timings(fields, h, w, "cpu", warmup, reps)
        dev = _render_timings(fields, h, w, gpu, 3, 10)
        out[name] = {
            "cpu_fwd_ms": round(cpu["fwd_s"] * 1e3, 3),
            "cpu_fwd_bwd_ms": round(cpu["fwd_bwd_s"] * 1e3, 3),
            "gpu_fwd_ms": round(dev["fwd_s"] * 1e3, 4),
            "gpu_fwd_bwd_ms": round(dev["fwd_bwd_s"] * 1e3, 4),
            "speedup_fwd": round(cpu["fwd_s"] / dev["fwd_s"], 1),
            "speedup_fwd_bwd": round(cpu["fwd_bwd_s"] / dev["fwd_bwd_s"], 1),
            "protocol": f"{warmup} warm-up(s), median of {reps}",
        }
    sec = out[f"{workload}_{W}x{H}"]["cpu_fwd_bwd_ms"] * 1e-3
    return {
        "value": round(bytes_step / sec / 1e9, 4),
        "unit": "GB/s",
        "cores": torch.get_num_threads(),
        "kind": "port",
        "sample": (f"the {workload} workload's fwd+bwd (whole field) through the torch "
                   "restatement of render.py (float32, render_chunk 512, autograd) on "
                   f"{threads} threads, 3 warm-ups and the median of 3: "
                   f"{sec * 1e3:.1f} ms per fwd+bwd; also the 256x256/2k CPU case"),
        "seconds_per_step": round(sec, 4),
        "os_cpu_count": os.cpu_count(),
        "cpu_model": _cpu_model(),
        "workloads": out,
        "gpu_timing": "render_hip_exact + autograd backward per call (binning, forward, "
                      "backward and the autograd node's allocations), 3 warm-ups, median of 10",
    }
def batched_fit(steps: int, num_fields: int = 64, seed: int = 0, plain_loop: bool = True) -> dict:
    """Batched fitting: `num_fields` fields, sides uniform over {256, 512, 768, 1024}, Gaussian
    counts log-uniform in [1k, 60k], fit-target recipe; the batched graph-captured fit against
    the plain per-image loop, same steps, same initial fields (fields sorted by size).  Without
    `plain_loop` only the batched fit runs (for profiling it alone)."""
    import numpy as np
    import torch

    from structsplat_amd import batch as B
    from structsplat_amd import fit as F
    from structsplat_amd import render as HR
    from structsplat_amd import synth
    from structsplat_amd import tune
    from structsplat_amd.gaussians import GaussianField

    dev = torch.device("cuda:0")
    sides, counts, targets, inits = synth.batched_fit_workload(num_fields, seed, device=dev)

    def fields():
        return [GaussianField.from_numpy(f["means"], f["scales"], f["angles"], f["colors"],
                                         opacities=f["opacity"], device=dev) for f in inits]

    cfg = F.FitConfig(iters=steps)
    batched = B.fit_batch(fields(), targets, cfg, graph=True)
    if not plain_loop:
        return {
            "fields": num_fields,
            "steps_per_field": steps,
            "batched_steps_per_s": round(steps / batched.seconds, 3),
            "batched_replay_ms_per_step": round(batched.step_seconds * 1e3
                                                / max(batched.timed_steps, 1), 4),
            "mean_psnr_db": round(float(np.mean(batched.psnr)), 3),
            "tile_size": HR.tile_edge(batched.tile_size),
            "schedule": "persistent" if batched.tile_size & HR.TILE_PERSISTENT else "grid",
        }
    # The plain loop at the batched fit's (autotuned) tile size: the same arithmetic per field.
    plain = B.fit_each(fields(), targets, dataclasses.replace(cfg, tile_size=batched.tile_size))
    diff = float(np.max(np.abs(np.array(batched.psnr) - np.array(plain.psnr))))
    return {
        "fields": num_fields,
        "steps_per_field": steps,
        "total_gaussians": int(counts.sum()),
        "total_pixels": int((sides.astype(np.int64) ** 2).sum()),
        "batched_steps_per_s": round(steps / batched.seconds, 3),
        "plain_steps_per_s": round(steps / plain.seconds, 3),
        "speedup": round(plain.seconds / batched.seconds, 3),
        "batched_replay_ms_per_step": round(batched.step_seconds * 1e3
                                            / max(batched.timed_steps, 1), 4),
        "plain_ms_per_step": round(plain.step_seconds * 1e3
                                   / max(plain.timed_steps // num_fields, 1), 4),
        "psnr_max_abs_diff_db": round(diff, 5),
        "mean_psnr_db": round(float(np.mean(batched.psnr)), 3),
        "tile_size": HR.tile_edge(batched.tile_size),
        "schedule": "persistent" if batched.tile_size & HR.TILE_PERSISTENT else "grid",
        "autotune_fwd_bwd_ms": ({tune.describe(k): round(v, 4) for k, v in batched.tune_ms.items()}
                                if batched.tune_ms else None),
        "note": "a step advances every field by one Adam step",
    }


def fit_loop(steps: int, seed: int = 0) -> dict:
    """The full fit loop: a 2048x2048 target (32 random sinusoids + 200 random step edges),
    200,000 Gaussians initialised uniformly (isotropic s = 0.5 sqrt(HW/N), angle 0, colors from
    the target, opacity 0.5), Adam (lr 1e-2 means, 5e-3 others), L1 + 0.2 (1 - SSIM).  The
    graph-captured step against the eager per-step loop, same steps, same initial field."""
    import torch

    from structsplat_amd import fit as F
    from structsplat_amd import synth
    from structsplat_amd.gaussians import GaussianField

    dev = torch.device("cuda:0")
    target = synth.fit_target(seed, 2048, 2048, device=dev)
    init = synth.fit_init_field(target, 200_000, seed)

    def field():
        return GaussianField.from_numpy(init["means"], init["scales"], init["angles"],
                                        init["colors"], opacities=init["opacity"], device=dev)

    cfg = F.FitConfig(iters=steps)
    graphed = F.fit(field(), target, cfg, graph=True)
    eager = F.fit(field(), target, cfg, graph=False)
    return {
        "steps": steps,
        "graph_steps_per_s": round(steps / graphed["seconds"], 3),
        "graph_ms_per_step": round(graphed["seconds"] * 1e3 / steps, 4),
        "eager_steps_per_s": round(steps / eager["seconds"], 3),
        "eager_ms_per_step": round(eager["seconds"] * 1e3 / steps, 4),
        "graph_replay_ms_per_step": round(graphed["step_seconds"] * 1e3
                                          / max(graphed["timed_steps"], 1), 4),
        "eager_step_ms": round(eager["step_seconds"] * 1e3 / max(eager["timed_steps"], 1), 4),
        "graph_psnr_db": round(graphed["psnr"], 4),
        "eager_psnr_db": round(eager["psnr"], 4),
        "tile_size": graphed["tile_size"],
        "regrows": graphed["regrows"],
        "note": ("steps_per_s and ms_per_step: the whole loop (capture, windows' snapshots and "
                 "readbacks); graph_replay_ms_per_step / eager_step_ms: the steps alone"),
    }


def band_shard(means, conics, colors, radii, op, grad_out, H: int, W: int, ts: int, steps: int,
               warmup: int, bands: int) -> dict:
    """Each row band of the image as one GPU's share of a band-split run: the band's members
    (its Gaussians plus the halo from neighbouring bands) binned, rendered and differentiated
    as an image of their own, with inputs resident.  The slowest band sets a multi-GPU step's
    pace; the all-reduce of the partial gradients is not part of this."""
    import torch

    from structsplat_amd import render as HR
    from structsplat_amd import shard

    dev = means.device
    per_band = []
    for y0, y1 in shard.band_bounds(H, bands, 32):
        _, m, q, c, r, o = shard.band_inputs(means, conics, colors, radii, op, y0, y1)
        m, q, c, r = (t.contiguous() for t in (m, q, c, r))
        o = m.new_empty(0) if o is None else o.contiguous()
        hb, n = y1 - y0, m.shape[0]
        _, off0 = HR.build_tile_index(m, q, r, hb, W, ts)
        P = int(off0[-1].item())
        cap = P + P // 8 + 1024
        gids = torch.empty(cap, dtype=torch.int32, device=dev)
        offsets = torch.empty(HR.num_tiles(hb, W, ts) + 1, dtype=torch.int32, device=dev)
        status = torch.zeros(1, dtype=torch.int32, device=dev)
        ws = HR.workspace(n, hb, W, ts, dev)
        slots = HR.grad_slots(cap, dev)
        img = torch.empty(hb, W, 3, device=dev)
        g = grad_out[y0:y1].contiguous()
        grads = (torch.empty_like(m), torch.empty_like(q), torch.empty_like(c),
                 torch.empty_like(o))

        def step():
            HR.build_tile_index_capacity(m, q, r, hb, W, cap, ts, out=(gids, offsets, status, ws))
            HR.forward_tiled(m, q, c, r, o, gids, offsets, hb, W, True, 1e-8, ts, False, 3.0,
                             out=img, ws=ws)
            HR.backward_tiled(g, m, q, c, r, o, None, img, gids, offsets, hb, W, True, 1e-8, ts,
                              False, 3.0, grads=grads, ws=ws, slots=slots, slot_capacity=cap,
                              status=status, after_forward=True)

        for _ in range(warmup):
            step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            step()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / steps
        if int(status.item()) != 0:
            raise RuntimeError("pair capacity exceeded in a band")
        per_band.append({"rows": [y0, y1], "gaussians": n, "tile_pairs": P,
                         "ms_per_step": round(ms, 4)})
    return {"bands": bands, "tile_size": HR.tile_edge(ts),
            "max_band_ms_per_step": max(b["ms_per_step"] for b in per_band),
            "member_gaussians_over_n": round(sum(b["gaussians"] for b in per_band)
                                             / means.shape[0], 4),
            "per_band": per_band,
            "note": "one GPU's share of a row-band split, timed band by band on one GPU; "
                    "the gradient all-reduce is not included"}


def copy_probe(mib: int) -> int:
    """One device-to-device copy of `mib` MiB (the calibration kernel of the HBM counters)."""
    import torch

    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": n * 4, "copy_probe_bytes_written": n * 4}))
    return 0


def valu_block(E: int, ms_per_step: float, kernel_ms: dict, stored: bool) -> dict:
    """The compute-side roof next to the HBM one: useful FP32 lane-operations (E visits times
    VISIT_OPS) per second against the vector issue rate."""
    bwd_ops = VISIT_OPS["bwd_stored" if stored else "bwd_recompute"]
    step_ops = E * (VISIT_OPS["fwd"] + bwd_ops)
    bwd = E * bwd_ops / (kernel_ms["k_backward"] * 1e-3)
    fwd = E * VISIT_OPS["fwd"] / (kernel_ms["k_forward"] * 1e-3)
    return {
        "visits": E,
        "ops_per_visit": {"fwd": VISIT_OPS["fwd"], "bwd": bwd_ops},
        "peak_lane_ops_per_s": VALU_PEAK_LANE_OPS,
        "step_frac": round(step_ops / (ms_per_step * 1e-3) / VALU_PEAK_LANE_OPS, 4),
        "k_forward_frac": round(fwd / VALU_PEAK_LANE_OPS, 4),
        "k_backward_frac": round(bwd / VALU_PEAK_LANE_OPS, 4),
        "note": ("useful lane-operations only: lanes outside a Gaussian's rectangle, reductions, "
                 "staging and addressing are not counted"),
    }


def measured_traffic(workload: str, kernel: str):
    """HBM bytes per launch of `kernel` from profiles/hbm_traffic.json, if it covers `workload`."""
    try:
        with open(TRAFFIC_FILE) as fh:
            data = json.load(fh)
    except (OSError, ValueError):
        return None
    entry = data.get("workloads", {}).get(workload, {}).get("kernels", {}).get(kernel)
    return entry.get("hbm_bytes_per_launch") if entry else None


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def _progress(msg: str) -> None:
    """One line on stderr when a phase of the run ends (the JSON line goes to stdout)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # 100 untimed steps bring the host's launch path and the GPU's clocks to a steady state (with
    # 5 the eager step read about 10% slower at identical kernel times); 100 timed steps are
    # about 35 ms of work at 2048x2048/200k.
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--workload", default="metric", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--normalizer", default="recompute", choices=["recompute", "stored"],
                    help="recompute: the backward recomputes den (the default); stored: the "
                         "forward saves den for the backward (the reference's data flow)")
    ap.add_argument("--tile-size", default="auto", choices=["auto", "16", "32", "272", "288"],
                    help="render tile edge (272 / 288: 16 / 32 with the persistent schedule); "
                         "auto = the autotuner's choice on this field")
    ap.add_argument("--batched-steps", type=int, default=500,
                    help="Adam steps per field of the batched-vs-plain fit (0 skips it)")
    ap.add_argument("--fit-steps", type=int, default=1000,
                    help="Adam steps of the 2048x2048/200k fit loop (0 skips it)")
    ap.add_argument("--batched-only", action="store_true",
                    help="only run the batched-vs-plain fit (--batched-steps steps)")
    ap.add_argument("--no-plain-loop", action="store_true",
                    help="with --batched-only: the batched fit alone, without the plain loop")
    ap.add_argument("--bands", type=int, default=8,
                    help="row bands timed as one GPU's share each (0 skips)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run one device copy of MIB MiB (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU (a single MI355X); the band_shard entry of its line "
              "times each GPU's share of a row-band split. Run with --gpus 1.", file=sys.stderr)
        return 2
    if args.copy_probe > 0:
        return copy_probe(args.copy_probe)
    if args.batched_only:
        print(json.dumps({"batched": batched_fit(max(args.batched_steps, 1),
                                                 plain_loop=not args.no_plain_loop)}), flush=True)
        return 0

    import torch

    from structsplat_amd import render as HR
    from structsplat_amd import tune

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        return 2
    desc, H, W, _ = WORKLOADS[args.workload]
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    means, conics, colors, radii, op = make_field(args.workload, dev)
    N = means.shape[0]
    # Tile shape: the autotuner's choice (fwd+bwd of this field timed at each candidate).
    best, tile_ms = tune.tune_tile_size([(H, W)], [N], means, conics, colors, radii, op,
                                        normalize=True, eps=1e-8, support_fade=False,
                                        sigma_cutoff=3.0,
                                        recompute_normalizer=args.normalizer == "recompute")
    ts = best if args.tile_size == "auto" else int(args.tile_size)
    edge = HR.tile_edge(ts)  # the tile edge of the chosen configuration (ts may add the schedule)
    T = HR.num_tiles(H, W, ts)
    T16 = HR.num_tiles(H, W, 16)
    gen = torch.Generator(device=dev).manual_seed(0)
    grad_out = (torch.rand(H, W, 3, device=dev, generator=gen) - 0.5) / (3 * H * W)

    # Pair count at the reference's 16-pixel tiling (the byte model's P).
    _, offsets16 = HR.build_tile_index(means, conics, radii, H, W, 16)
    P16 = int(offsets16[-1].item())
    # Size the pair buffer once (outside the timed region), then bin without host syncs.
    _, offsets0 = HR.build_tile_index(means, conics, radii, H, W, ts)
    P = int(offsets0[-1].item())
    capacity = P + P // 8 + 1024
    gids = torch.empty(capacity, dtype=torch.int32, device=dev)
    offsets = torch.empty(T + 1, dtype=torch.int32, device=dev)
    status = torch.zeros(1, dtype=torch.int32, device=dev)
    ws = HR.workspace(N, H, W, ts, dev)
    slots = HR.grad_slots(capacity, dev)  # deterministic backward: one slot per pair
    out = torch.empty(H, W, 3, device=dev)
    den = torch.empty(H, W, device=dev)
    grads = (torch.empty_like(means), torch.empty_like(conics), torch.empty_like(colors),
             torch.empty_like(op))

    stream = torch.cuda.current_stream(dev)
    phase_ms = {"bin": [], "fwd": [], "bwd": []}

    stored = args.normalizer == "stored"

    def step(ev: dict | None, saved: bool = stored):
        """One fwd+bwd; `ev`: event pairs to record around its three phases; `saved`: the
        forward stores den and the backward reads it."""
        record = ev is not None
        if record:
            ev["b"][0].record(stream)
        HR.build_tile_index_capacity(means, conics, radii, H, W, capacity, ts,
                                     out=(gids, offsets, status, ws))
        if record:
            ev["b"][1].record(stream)
            ev["f"][0].record(stream)
        HR.forward_tiled(means, conics, colors, radii, op, gids, offsets, H, W, True, 1e-8, ts,
                         False, 3.0, out=out, den=den if saved else None, ws=ws)
        if record:
            ev["f"][1].record(stream)
            ev["g"][0].record(stream)
        HR.backward_tiled(grad_out, means, conics, colors, radii, op, den if saved else None, out,
                          gids, offsets, H, W, True, 1e-8, ts, False, 3.0, grads=grads, ws=ws,
                          slots=slots, slot_capacity=capacity, status=status, after_forward=True)
        if record:
            ev["g"][1].record(stream)

    for _ in range(args.warmup):
        step(None)
    torch.cuda.synchronize()
    if int(status.item()) != 0:
        raise RuntimeError("pair capacity exceeded during warmup")

    # Whole-step timing: K back-to-back steps bracketed by device synchronization.
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(None)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / args.steps

    # The same step captured once as a HIP graph and replayed (how the fit loop runs it): the
    # kernels are identical, the launch gaps between them shrink.
    side = torch.cuda.Stream(device=dev)
    side.wait_stream(stream)
    with torch.cuda.stream(side):
        step(None)  # warm the side stream once before capture
    stream.wait_stream(side)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        step(None)
    for _ in range(args.warmup):
        graph.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        graph.replay()
    torch.cuda.synchronize()
    graph_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    if int(status.item()) != 0:
        raise RuntimeError("pair capacity exceeded in the graph replays")

    # Per-phase timing with events on the launch stream: a separate pass of K back-to-back
    # steps, each phase between its own event pair, one synchronization at the end. The device
    # stays ahead of the host, so a phase's time is the device's (synchronizing after every
    # step would add the host's launch latency to each step's first phase).
    evs = [{k: [torch.cuda.Event(enable_timing=True) for _ in range(2)] for k in ("b", "f", "g")}
           for _ in range(args.steps)]
    for ev in evs:
        step(ev)
    torch.cuda.synchronize()
    for ev in evs:
        phase_ms["bin"].append(ev["b"][0].elapsed_time(ev["b"][1]))
        phase_ms["fwd"].append(ev["f"][0].elapsed_time(ev["f"][1]))
        phase_ms["bwd"].append(ev["g"][0].elapsed_time(ev["g"][1]))
    if int(status.item()) != 0:
        raise RuntimeError("pair capacity exceeded")
    phase = {k: statistics.mean(v) for k, v in phase_ms.items()}

    # Kernel-only durations: the library's event pairs around k_forward / k_backward.
    from structsplat_amd import _native

    def kernel_times(saved: bool) -> dict[str, float]:
        _native.kernel_timing_read(_native.KERNEL_FORWARD)
        _native.kernel_timing_read(_native.KERNEL_BACKWARD)
        _native.kernel_timing(True)
        try:
            for _ in range(args.steps):
                step(None, saved)
            torch.cuda.synchronize()
        finally:
            _native.kernel_timing(False)
        nf, fwd_kernel_ms = _native.kernel_timing_read(_native.KERNEL_FORWARD)
        nb, bwd_kernel_ms = _native.kernel_timing_read(_native.KERNEL_BACKWARD)
        return {"k_forward": fwd_kernel_ms / max(nf, 1),
                "k_backward": bwd_kernel_ms / max(nb, 1)}

    kernel_ms = kernel_times(stored)
    # The other normalizer mode: whole-step wall time and kernel times.
    for _ in range(args.warmup):
        step(None, not stored)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(None, not stored)
    torch.cuda.synchronize()
    other_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    other_kernel_ms = kernel_times(not stored)

    E = rect_visits(means, radii, H, W)
    P_now = int(offsets[-1].item())
    if P_now != P:
        raise RuntimeError(f"pair count changed between binnings: {P} -> {P_now}")
    b = algorithmic_bytes(N, P16, T16, H * W)       # the workload's work (reference tiling)
    b_tile = algorithmic_bytes(N, P, T, H * W)      # what the chosen tiling must move
    value = b["step"] / (ms_per_step * 1e-3) / 1e9
    bwd_bytes = backward_kernel_bytes(P, H * W, stored)
    bwd_gbs = bwd_bytes / (kernel_ms["k_backward"] * 1e-3) / 1e9
    bwd_kernel = (f"k_backward_persistent<{edge}>" if ts & HR.TILE_PERSISTENT
                  else f"k_backward<{edge}>") + ("/stored" if stored else "")
    line = {
        "metric": "fwd+bwd ms & HBM GB/s at 2048^2/200k Gaussians; batched fit steps/s vs plain",
        "value": round(value, 3),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": desc,
            "H": H, "W": W, "n_gaussians": N,
            "tile_size": edge, "schedule": ("persistent" if ts & HR.TILE_PERSISTENT
                                            else "grid"),
            "tile_choice": args.tile_size,
            "autotune_fwd_bwd_ms": {tune.describe(k): round(v, 4) for k, v in tile_ms.items()},
            "tile_pairs": P, "tiles": T, "tile_pairs_at_16": P16,
            "phase_ms": {k: round(v, 4) for k, v in phase.items()},
            "kernel_ms": {k: round(v, 4) for k, v in kernel_ms.items()},
            "normalizer": ("stored by the forward, read by the backward" if stored
                           else "recomputed in the backward (den not stored)"),
            "graph_replay_ms_per_step": round(graph_ms, 4),
            "timing": ("ms_per_step and value: K eager steps (host launches) between two "
                       "synchronizations; graph_replay_ms_per_step: the same step captured "
                       "once as a HIP graph and replayed K times"),
            ("recomputed_normalizer_variant" if stored else "stored_normalizer_variant"): {
                "ms_per_step": round(other_ms, 4),
                "kernel_ms": {k: round(v, 4) for k, v in other_kernel_ms.items()},
            },
            "algorithmic_bytes": b,
            "algorithmic_bytes_at_tile": b_tile,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": bwd_kernel,
            "algorithmic_bytes_per_launch": bwd_bytes,
            "achieved": round(bwd_gbs, 3),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(bwd_gbs / HBM_PEAK_GBS, 5),
            "traffic": measured_traffic(args.workload, bwd_kernel),
        },
        "valu": valu_block(E, ms_per_step, kernel_ms, stored),
    }
    _progress(f"fwd+bwd {ms_per_step:.4f} ms per step")
    if args.bands > 0:
        line["band_shard"] = band_shard(means, conics, colors, radii, op, grad_out, H, W, ts,
                                        args.steps, args.warmup, args.bands)
        _progress("band shares timed")
    if args.batched_steps > 0:
        line["batched"] = batched_fit(args.batched_steps)
        _progress(f"batched fit: {line['batched']['speedup']}x the plain loop")
    if args.fit_steps > 0:
        line["fit"] = fit_loop(args.fit_steps)
        _progress(f"fit loop: {line['fit']['graph_ms_per_step']} ms per graphed step")
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(args.workload, H, W, b["step"])
        _progress("CPU baseline timed")
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
