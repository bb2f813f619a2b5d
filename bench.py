"""Benchmark: multislice mixed-probe ptychography forward + adjoint throughput on one MI355X.

Workload (8 slices of 384x384, 2 A thick; 128x128 patterns; 4 probe modes; 64x64 raster scan at a
4 px step with +-0.5 px jitter = 4096 patterns; 300 kV; Poisson data at 1e5 e-/pattern; complex64;
batch 256).  One step = one pass over all 4096 patterns in 16 batches, each batch running what the
reference's reconstruction loop runs per batch (quantem ptychography.py:213-243): fused forward +
loss, hand-written adjoint (object, probe and scan-position gradients), the object-gradient
rescale and an Adam step.  Inputs are resident in HBM before the timed region.

Prints one JSON line.  With --gpus N > 1 it exits: this build measures one GPU.
"""

from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np
import torch

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from quantem_amd import _native  # noqa: E402
from quantem_amd import ptychography as qp  # noqa: E402
from quantem_amd import optim as qoptim  # noqa: E402
from quantem_amd import synth  # noqa: E402

METRIC = "patterns/s fwd+adjoint, % HBM roofline; time-to-target-NMSE warm-start vs flat"
FP32_PEAK_TFLOPS = 157.3  # MI355X vector FP32 with packed (2-wide) FMA
FP32_UNPACKED_TFLOPS = 78.6  # the same without packed FP32, which this library's device code is built without
HBM_PEAK_GBS = 8000.0
CALIBRATION_BYTES = 1 << 30  # torch.neg over 1 GiB: 1 GiB read, 1 GiB written


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def fft_flops(n) -> float:
    """Textbook complex 2-D FFT flop count 5 P log2(P) for P = Nr * Nc pixels (n = N or (Nr, Nc))."""
    nr, nc = synth._shape2(n)
    return 5.0 * nr * nc * np.log2(nr * nc)


def make_inputs(spec: synth.ProblemSpec, dev: torch.device, seed: int = 0) -> dict:
    prob = synth.make_problem(spec, seed=seed)
    obj_true = torch.from_numpy(prob["obj"]).to(dev)
    probe = torch.from_numpy(prob["probe"]).to(dev)
    pos = torch.from_numpy(prob["positions"]).to(dev)
    props = torch.from_numpy(prob["propagators"]).to(dev)
    P = pos.shape[0]
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed + 1)
    amps = torch.empty((P,) + synth._shape2(spec.roi), dtype=torch.float32, device=dev)
    total = 0.0
    for i in range(0, P, spec.batch):
        I = qp.forward_intensities(obj_true, probe, pos[i:i + spec.batch], props)
        counts = torch.poisson(I.clamp_min(0), generator=gen)
        total += float(counts.sum().item())
        amps[i:i + spec.batch] = torch.sqrt(counts)
    return {"obj_true": obj_true, "probe": probe, "positions": pos, "propagators": props, "targets": amps,
            "mean_intensity": total / P, "num_gpts": P}


def batches(P: int, B: int, seed: int) -> list:
    perm = np.random.default_rng(seed).permutation(P)
    return [torch.from_numpy(perm[i:i + B].copy()) for i in range(0, P, B)]



def gpu_pass(state: dict, data: dict, idx_list: list, segment_slices: int = 0) -> None:
    obj, probe, opt = state["obj"], state["probe"], state["opt"]
    for idx in idx_list:
        opt.zero_grad(set_to_none=True)
        loss = qp.multislice_loss(obj, probe, data["positions"][idx], data["propagators"], data["targets"][idx],
                                  None, None, num_gpts=data["num_gpts"], mean_intensity=data["mean_intensity"],
                                  segment_slices=segment_slices)
        loss.backward()
        opt.step()  # object-gradient rescale (grad_scale of the object group) + Adam, one kernel per tensor


def host_cpu() -> dict:
    """Model name, physical cores and hardware threads of the host (from /proc/cpuinfo), and the
    threads this process may run on."""
    model, cores, threads = platform.processor() or platform.machine(), set(), 0
    try:
        phys = core = None
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "processor":
                    threads += 1
                elif k == "model name":
                    model = v
                elif k == "physical id":
                    phys = v
                elif k == "core id":
                    core = v
                    cores.add((phys, core))
    except OSError:
        pass
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return {"model": model, "physical_cores": len(cores) or None, "hw_threads": threads or None,
            "usable_threads": usable}


def cpu_baseline(data: dict, spec: synth.ProblemSpec, seconds: float, sample_batch: int) -> dict:
    """The reference op order on the host cores (complex64 torch, loss.backward()), bounded sample."""
    from oracle import torch_ref

    host = host_cpu()
    threads = host["usable_threads"]
    torch.set_num_threads(threads)
    idx = torch.arange(sample_batch)
    obj = torch.ones_like(data["obj_true"]).cpu().requires_grad_(True)
    probe = data["probe"].cpu().clone().requires_grad_(True)
    pos = data["positions"][idx].cpu()
    props = data["propagators"].cpu()
    tgt = data["targets"][idx.to(data["targets"].device)].cpu()

    def one():
        obj.grad = None
        probe.grad = None
        loss, _ = torch_ref.batch_loss(obj, probe, pos, props, tgt, None, data["num_gpts"], data["mean_intensity"])
        loss.backward()

    one()  # warm-up
    n = 0
    t0 = time.perf_counter()
    while True:
        one()
        n += 1
        el = time.perf_counter() - t0
        if el >= seconds:
            break
    return {"value": n * sample_batch / el, "unit": "patterns/s", "cores": torch.get_num_threads(), "kind": "port",
            "sample": f"{n} batches x {sample_batch} patterns of the same workload (torch {torch.__version__} "
                      f"complex64 CPU, reference op order, loss.backward()), {el:.1f} s, "
                      f"{torch.get_num_threads()} threads on {host['model']} ({host['physical_cores']} physical cores, "
                      f"{host['hw_threads']} hardware threads on the host)",
            "host": host}


def counter_child(batches_to_run: int = 2) -> int:
    """Profiled child of measure_traffic(): torch.neg over 1 GiB (the calibration kernel), then
    `batches_to_run` forward + loss + backward batches of the headline workload's shapes."""
    dev = torch.device("cuda:0")
    _native.load()
    src = torch.ones(CALIBRATION_BYTES // 4, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    torch.neg(src, out=dst)
    torch.cuda.synchronize()
    del src, dst
    spec = synth.HEADLINE
    S, (H, W), N, M, B = spec.num_slices, spec.obj_shape, spec.roi, spec.num_modes, spec.batch
    prob = synth.make_problem(spec, seed=0)
    obj = torch.from_numpy(prob["obj"]).to(dev).requires_grad_(True)
    probe = torch.from_numpy(prob["probe"]).to(dev).requires_grad_(True)
    pos = torch.from_numpy(prob["positions"][:B]).to(dev)
    props = torch.from_numpy(prob["propagators"]).to(dev)
    tg = torch.rand((B, N, N), device=dev) * 10
    for _ in range(batches_to_run):
        loss = qp.multislice_loss(obj, probe, pos, props, tg, None, None, num_gpts=4096, mean_intensity=1e5)
        loss.backward()
    torch.cuda.synchronize()
    return 0


def measure_traffic(timeout_s: float = 150.0) -> dict | None:
    """HBM bytes per dispatch of the multislice kernels at the headline shape, measured in this run:
    two rocprofv3 passes (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch) over a child process
    (counter_child), each scaled by what the same counter reads for the 1 GiB calibration copy."""
    import csv
    import glob
    import shutil
    import subprocess
    import tempfile

    prof = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not os.path.exists(prof):
        return None
    vals = {}
    with tempfile.TemporaryDirectory(prefix="qem_counters_") as tmp:
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            out = os.path.join(tmp, counter)
            cmd = [prof, "--pmc", counter, "--output-format", "csv", "-d", out, "-o", "run", "--",
                   sys.executable, os.path.abspath(__file__), "--counter-child"]
            try:
                r = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout_s, start_new_session=True,
                                   cwd=tmp)
            except subprocess.TimeoutExpired:
                log(f"counter pass {counter} timed out")
                return None
            files = glob.glob(os.path.join(out, "**", "*counter_collection.csv"), recursive=True)
            if not files:
                log(f"counter pass {counter}: no counter file (rc {r.returncode}): {r.stderr[-500:]}")
                return None
            per = {}
            with open(files[0]) as f:
                for row in csv.DictReader(f):
                    if row.get("Counter_Name") != counter:
                        continue
                    name = row.get("Kernel_Name", "")
                    per.setdefault(name, []).append(float(row["Counter_Value"]))
            vals[counter] = per

    def pick(per, *keys):
        for name, v in per.items():
            if all(k in name for k in keys):
                return float(np.mean(v)), len(v)
        return None, 0

    cal_f, _ = pick(vals["FETCH_SIZE"], "neg")
    cal_w, _ = pick(vals["WRITE_SIZE"], "neg")
    if not cal_f or not cal_w:
        return None
    sf = CALIBRATION_BYTES / (cal_f * 1024.0)
    sw = CALIBRATION_BYTES / (cal_w * 1024.0)
    kernels = {"adjoint": ("k_adjoint<128",), "forward": ("k_forward<128, false, true",),
               "object_reduce": ("k_obj_reduce",), "loss": ("k_loss<",), "probe_reduce": ("k_probe_reduce",)}
    res = {"scale_fetch": round(sf, 4), "scale_write": round(sw, 4), "kernels": {}}
    for k, keys in kernels.items():
        f, nf = pick(vals["FETCH_SIZE"], *keys)
        w, nw = pick(vals["WRITE_SIZE"], *keys)
        if f is None or w is None:
            continue
        res["kernels"][k] = {"read": f * 1024.0 * sf, "write": w * 1024.0 * sw, "dispatches": nf}
    return res


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-batch", type=int, default=256, help="patterns per CPU batch (the GPU batch)")
    ap.add_argument("--no-counters", action="store_true", help="skip the two rocprofv3 counter passes")
    ap.add_argument("--counter-child", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--workload", default="headline", choices=sorted(synth.WORKLOADS),
                    help="headline: 8 x 384^2, 128^2, 4 modes (the bench line); depth20-256px: 20 x 512^2, "
                         "256^2, 2 modes; large30-1024px: 30 x 1024^2, 256^2, 8 modes, 16384 patterns; "
                         "single-slice: the reference's CPU case")
    ap.add_argument("--segment", type=int, default=0,
                    help="segment_slices of the backward pass: 0 stores every incident wave, k >= 2 keeps every "
                         "k-th and recomputes the rest")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run one elementwise kernel over 1 GiB first (torch.neg: 1 GiB read, 1 GiB written), "
                         "the reference point for the HBM byte counters")
    args = ap.parse_args()
    if args.counter_child:
        return counter_child()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (run with --gpus 1)",
              file=sys.stderr, flush=True)
        return 2
    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        return 1
    dev = torch.device("cuda:0")
    _native.load()
    spec = synth.WORKLOADS[args.workload]
    if args.calibrate_copy:
        src = torch.ones(1 << 28, dtype=torch.float32, device=dev)  # 1 GiB
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        torch.neg(src, out=dst)
        torch.cuda.synchronize()
        log("calibration kernel torch.neg: 1073741824 bytes read, 1073741824 bytes written")
        del src, dst
    log(f"generating workload {spec.name}")
    data = make_inputs(spec, dev, args.seed)
    S, (H, W), N, M, B = spec.num_slices, spec.obj_shape, spec.roi, spec.num_modes, spec.batch
    P = data["num_gpts"]
    obj = torch.nn.Parameter(torch.ones((S, H, W), dtype=torch.complex64, device=dev))
    probe = torch.nn.Parameter(data["probe"].clone())
    # the reference's object-gradient rescale upsample_factor**2 / 2 (ptychography.py:330-332,
    # dataset_models.py:451-456, 877-879): 252 / 64 = 3.9375 at this workload
    scale = synth.spec_upsample_factor(spec) ** 2 / 2
    opt = qoptim.Adam([{"params": [obj], "lr": 5e-3, "grad_scale": scale}, {"params": [probe], "lr": 1e-3}])
    state = {"obj": obj, "probe": probe, "opt": opt}
    idx_list = [i.to(dev) for i in batches(P, B, args.seed)]

    log(f"warmup {args.warmup} steps")
    for _ in range(args.warmup):
        gpu_pass(state, data, idx_list, args.segment)
    torch.cuda.synchronize()
    torch.cuda.reset_peak_memory_stats(dev)
    log(f"timing {args.steps} steps")
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    with _native.KernelTimer(max_launches=len(idx_list) * args.steps + 8) as kt:
        torch.cuda.synchronize()
        ev0.record()
        for _ in range(args.steps):
            gpu_pass(state, data, idx_list, args.segment)
        ev1.record()
        torch.cuda.synchronize()
    total_ms = ev0.elapsed_time(ev1)
    peak_hbm = torch.cuda.max_memory_allocated(dev)
    ms_per_step = total_ms / args.steps
    value = P * args.steps / (total_ms / 1e3)
    log(f"{ms_per_step:.3f} ms/step, {value:.0f} patterns/s; fwd kernel {kt.fwd_ms / max(kt.n_fwd, 1):.3f} ms x "
        f"{kt.n_fwd}, adjoint kernel {kt.adj_ms / max(kt.n_adj, 1):.3f} ms x {kt.n_adj}")

    # roofline of the dominant kernel: FFT flops, 5 N^2 log2(N^2) per 2-D transform, 2S transforms per (pattern, mode)
    flops_per_launch = B * M * 2 * S * fft_flops(N)
    t_fwd = kt.fwd_ms / max(kt.n_fwd, 1)
    t_adj = kt.adj_ms / max(kt.n_adj, 1)
    dom, t_dom = ("multislice adjoint kernel", t_adj) if t_adj >= t_fwd else ("multislice forward kernel", t_fwd)
    achieved = flops_per_launch / (t_dom / 1e3) / 1e12
    # HBM bytes per launch from the counter passes (headline shapes, every wave stored)
    counters = None
    if not args.no_counters and args.workload == "headline" and args.segment == 0:
        log("counter passes (rocprofv3 FETCH_SIZE, WRITE_SIZE)")
        counters = measure_traffic()
    key = "adjoint" if dom == "multislice adjoint kernel" else "forward"
    kc = (counters or {}).get("kernels", {}).get(key)
    traffic = (kc["read"] + kc["write"]) if kc else None
    roof = {"bound": "vector-fp32", "achieved": round(achieved, 3), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(achieved / FP32_PEAK_TFLOPS, 4),
            "peak_unpacked": FP32_UNPACKED_TFLOPS, "frac_unpacked": round(achieved / FP32_UNPACKED_TFLOPS, 4),
            "traffic": round(traffic) if traffic else None,
            "traffic_gbs": round(traffic / (t_dom / 1e3) / 1e9, 1) if traffic else None,
            "kernel": dom, "kernel_ms": round(t_dom, 4),
            "fwd_kernel_ms": round(t_fwd, 4), "adj_kernel_ms": round(t_adj, 4),
            "note": "FFT-flop bound on the FP32 vector units (no dense contraction, no MFMA); peak 157.3 TFLOP/s "
                    "assumes packed FP32, which this build disables, so 78.6 TFLOP/s is the reachable vector peak "
                    "(frac_unpacked); algorithmic flops = B*M*2S*5N^2*log2(N^2) per launch; traffic = HBM bytes "
                    "per launch of that kernel from rocprofv3 FETCH_SIZE / WRITE_SIZE measured in this run, each "
                    "scaled by the counter's reading for a 1 GiB copy"}
    if counters:
        alg_batch = B * (24 * S * N * N + 4 * N * N) + 16 * M * N * N + 8 * (S - 1) * N * N + 8 * S * H * W
        ks = counters["kernels"]
        per_batch = sum(v["read"] + v["write"] for v in ks.values())
        roof["counters"] = {k: {"read_GB": round(v["read"] / 1e9, 4), "write_GB": round(v["write"] / 1e9, 4)}
                            for k, v in ks.items()}
        roof["counter_scale"] = {"fetch": counters["scale_fetch"], "write": counters["scale_write"]}
        roof["batch_traffic_GB"] = round(per_batch / 1e9, 3)
        roof["batch_algorithmic_GB"] = round(alg_batch / 1e9, 3)
    cpu = None
    if not args.no_cpu:
        log(f"cpu baseline ~{args.cpu_seconds:.0f} s")
        cpu = cpu_baseline(data, spec, args.cpu_seconds, args.cpu_batch)
        log(f"cpu {cpu['value']:.1f} patterns/s")
    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "patterns/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "complex64",
        "data": "synthetic",
        "config": {"workload": spec.name, "slices": S, "object": [H, W], "roi": N if isinstance(N, int) else list(N),
                   "modes": M,
                   "scan": list(spec.scan), "patterns": P, "batch": B, "parallelism": "single-gpu",
                   "segment_slices": args.segment},
        "peak_hbm_bytes": int(peak_hbm),
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
