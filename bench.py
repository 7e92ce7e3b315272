"""Benchmark of the spherical-harmonic transform on one MI355X.

A step is one forward+inverse pass of the transform over one batch: T85L32
(256 x 128 Gaussian grid, vorticity, divergence, temperature on 32 levels plus
ln p_s = 97 fields), spectral -> grid -> spectral, with inputs resident in HBM.
fields/s counts fields that finish the round trip.  Prints one JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision fp64|bf16x3|fp16x3|bf16x6|auto]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "dennou-ccm_amd", "py")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

METRIC = "T85L32 SHT fields/s, split-BF16 MFMA vs FP64 at <=1e-6 err; serve p50/p99 ms"
N, IM, JM, NLEV = 85, 256, 128, 32
NF = 3 * NLEV + 1
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
BF16_PEAK_TFLOPS = 2500.0

PRECISIONS = {"fp64": 0, "bf16x3": 1, "fp16x3": 2, "bf16x6": 3}
SPIN_CYCLES = 200_000_000  # ~0.1 s: longer than it takes the host to queue K steps


def algorithmic(plan_tables_bytes: dict, nf: int, four_bytes: int = 16):
    """Per-launch algorithmic bytes and flops of each stage (FP64 grid and
    spectra; the Fourier intermediate is complex FP64, or FP32 in mixed plans)."""
    K = (N + 1) * (N + 2) // 2
    four = four_bytes * (N + 1) * JM  # Fourier intermediate per field
    grid = 8 * IM * JM
    spec = 16 * K
    fft_flops = 2.5 * JM * IM * np.log2(IM)
    leg_flops = 2.0 * JM * K
    return {
        "leg_inv": (nf * (spec + four) + plan_tables_bytes["inv"], nf * leg_flops),
        "fft_inv": (nf * (four + grid), nf * fft_flops),
        "fft_fwd": (nf * (grid + four), nf * fft_flops),
        "leg_fwd": (nf * (four + spec) + plan_tables_bytes["fwd"], nf * leg_flops),
    }


def table_bytes(precision: int):
    """Bytes of the Legendre tables one launch streams (padded layouts)."""
    P2p = ((JM // 2) + 15) // 16 * 16
    P2p32 = (P2p + 31) // 32 * 32
    pieces = {0: 0, 1: 2, 2: 2, 3: 3}[precision]
    inv = fwd = 0
    for m in range(N + 1):
        ke = (N - m + 2) // 2
        if precision == 0:
            inv += 8 * 2 * ((ke + 3) // 4 * 4) * P2p
            fwd += 8 * 2 * P2p * ((ke + 15) // 16 * 16)
        else:
            inv += 2 * pieces * 2 * P2p * ((ke + 31) // 32 * 32)
            fwd += 2 * pieces * 2 * ((ke + 15) // 16 * 16) * P2p32
    return {"inv": inv, "fwd": fwd}


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a - b)) / np.linalg.norm(np.ravel(b)))


def run_gpu(precision: int, steps: int, warmup: int, autotune: bool = True, fused: bool = False):
    import torch

    import shtamd
    from shtamd import synth

    dev = torch.device("cuda:0")
    spec_c = synth.state_spectra(N, NLEV)
    plan = shtamd.Plan(N, IM, JM, NF, precision=precision)
    if autotune:
        shapes, _ = plan.autotune(NF, reps=20)  # Legendre kernel shapes for this batch
    else:
        if fused and precision != 0:  # the fused inverse of the mixed forms
            plan.set_shapes(0, shtamd.SHAPE_FUSED, shtamd.SHAPE_AUTO)
        shapes = plan.get_shapes(NF)  # the built-in choice
    s_in = torch.from_numpy(shtamd.spec_to_real(spec_c)).to(dev)
    grid = torch.empty((NF, JM, IM), dtype=torch.float64, device=dev)
    s_out = torch.empty_like(s_in)
    stream = torch.cuda.current_stream()

    def step():
        plan.spec_to_grid(s_in, grid, NF, stream)
        plan.grid_to_spec(grid, s_out, NF, stream)

    for _ in range(max(warmup, 1)):
        step()
    torch.cuda.synchronize()
    out_grid = grid.cpu().numpy().copy()
    out_spec = shtamd.real_to_spec(s_out.cpu().numpy())

    # Each timed loop starts behind a bounded spin kernel (torch.cuda._sleep), so
    # the host has queued all K steps before the GPU reaches them: the events
    # then bracket back-to-back GPU execution, not host launch latency.
    def prefill():
        torch.cuda._sleep(SPIN_CYCLES)

    # eager, K steps bracketed by events on the launch stream
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    prefill()
    e0.record(stream)
    for _ in range(steps):
        step()
    e1.record(stream)
    torch.cuda.synchronize()
    eager_ms = e0.elapsed_time(e1) / steps

    # the same step captured once as a HIP graph and replayed K times
    graph_ms = None
    try:
        g = torch.cuda.CUDAGraph()
        cap = torch.cuda.Stream()
        cap.wait_stream(stream)
        with torch.cuda.stream(cap):
            step()  # warm the capture stream
        torch.cuda.synchronize()
        with torch.cuda.graph(g, stream=cap):
            plan.spec_to_grid(s_in, grid, NF, cap)
            plan.grid_to_spec(grid, s_out, NF, cap)
        for _ in range(max(warmup, 1)):
            g.replay()
        torch.cuda.synchronize()
        if not torch.equal(s_out.cpu(), torch.from_numpy(shtamd.spec_to_real(out_spec))):
            raise RuntimeError("graph replay differs from eager")
        # replay() launches on the current stream: bracket it there
        g0, g1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        prefill()
        g0.record(stream)
        for _ in range(steps):
            g.replay()
        g1.record(stream)
        torch.cuda.synchronize()
        graph_ms = g0.elapsed_time(g1) / steps
    except Exception as exc:  # graph capture is an optimisation; report why it was skipped
        print(f"# graph replay skipped: {exc}", file=sys.stderr)

    # per-kernel durations: the library's own event timers around each launch
    plan.set_timing(True)
    prefill()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    st = plan.stage_times()
    plan.set_timing(False)
    kernel_ms = {k: (v[0] / v[1] if v[1] else None) for k, v in st.items()}
    plan.close()
    return {"eager_ms": eager_ms, "graph_ms": graph_ms, "kernel_ms": kernel_ms,
            "grid": out_grid, "spec": out_spec, "spec_in": spec_c, "shapes": shapes}


def host_cpu_info():
    """CPU model, sockets and logical CPUs of this host (from /proc/cpuinfo),
    and the CPUs this process may run on."""
    model, sockets, logical = None, set(), 0
    try:
        for line in open("/proc/cpuinfo"):
            key, _, val = line.partition(":")
            key, val = key.strip(), val.strip()
            if key == "model name" and model is None:
                model = val
            elif key == "physical id":
                sockets.add(val)
            elif key == "processor":
                logical += 1
    except OSError:
        pass
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return {"model": model, "sockets": len(sockets) or None, "logical_cpus": logical or None,
            "usable_cpus": usable}


def cpu_baseline(spec_c, seconds: float, threads: int):
    import oracle

    p = oracle.Plan(N, IM, JM)
    t0 = time.perf_counter()
    n = 0
    while True:
        g = p.spec_to_grid(spec_c, nthreads=threads)
        p.grid_to_spec(g, nthreads=threads)
        n += 1
        dt = time.perf_counter() - t0
        if dt >= seconds and n >= 3:
            break
    return n * NF / dt, n, dt


def run_serve(precision: str, rates, nreq_per_rate, slots: int = 4, max_fields: int = 256,
              err_sample: int = 100):
    """Serving workload: Poisson request streams; returns a JSON-ready dict."""
    import torch

    import oracle
    import shtamd
    from shtamd import serve

    grids = [(T, *serve.GRIDS[T]) for T in (21, 42, 85, 170)]
    out = {"precision": precision, "slots": slots, "max_fields": max_fields, "rates": {}}
    for rate in rates:
        n = int(nreq_per_rate(rate))
        reqs = serve.make_requests(n, seed=int(rate))
        data = serve.build(torch, reqs)
        srv = shtamd.Server(grids, slots=slots, max_fields=max_fields, precision=PRECISIONS[precision])
        lat, wall, retries = serve.run_stream(torch, srv, data, rate)
        srv.drain()
        st = srv.stats()
        srv.close()
        fields = sum(L for _, L, _ in reqs)
        # error against the FP64 oracle on a sample of the requests
        rng = np.random.default_rng(1)
        errs = []
        plans = {}
        for k in rng.choice(n, size=min(err_sample, n), replace=False):
            d = data[k]
            ref = plans.setdefault(d.T, oracle.Plan(d.T, *serve.GRIDS[d.T]))
            if d.direction == shtamd.DIR_INVERSE:
                errs.append(rel(d.out.cpu().numpy(), ref.spec_to_grid(d.spec)))
            else:
                errs.append(rel(shtamd.real_to_spec(d.out.cpu().numpy()), ref.grid_to_spec(d.inp.cpu().numpy())))
        out["rates"][str(rate)] = {
            "requests": n, "p50_ms": round(float(np.percentile(lat, 50)), 4),
            "p99_ms": round(float(np.percentile(lat, 99)), 4), "max_ms": round(float(np.max(lat)), 4),
            "fields_per_s": round(fields / wall, 1), "batches": st["batches"],
            "requests_per_batch": round(n / max(1, st["batches"]), 2), "refused_retries": retries,
            "max_rel_l2_vs_fp64_oracle": max(errs)}
        del data
        torch.cuda.empty_cache()
    return out


def run_host_dropin(precision: int, reps: int = 40):
    """The synchronous host entry points DCPAM's wa_module calls (xya_wa and
    wa_xya on host arrays in SPML packing; each call copies through pinned
    staging over PCIe and waits for the result): the 97-level T85L32 batch
    per call.  Returns ms per call and the round trip's fields/s."""
    import shtamd

    plan = shtamd.Plan(N, IM, JM, NF, precision=precision)
    rng = np.random.default_rng(5)
    wa = rng.standard_normal((NF, (N + 1) ** 2)) * 1e-2
    # the model's own arrays, reused call after call (a fresh 25 MB result
    # array per call adds its first-touch page faults)
    xya = plan.xya_wa(wa)
    wa_out = plan.wa_xya(xya)
    for _ in range(10):  # warm: staging, copy workers, the host's clocks after the GPU timing
        plan.xya_wa(wa, out=xya)
        plan.wa_xya(xya, out=wa_out)

    def per_call(fn):
        # median over the calls: a call that the host's scheduler delays
        # (the copies run on host threads) says little about the path
        t = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            t.append(time.perf_counter() - t0)
        return float(np.median(t)) * 1e3, float(np.mean(t)) * 1e3

    inv, inv_mean = per_call(lambda: plan.xya_wa(wa, out=xya))
    fwd, fwd_mean = per_call(lambda: plan.wa_xya(xya, out=wa_out))
    fresh, _ = per_call(lambda: plan.xya_wa(wa))
    plan.close()
    return {"xya_wa_ms": round(inv, 4), "wa_xya_ms": round(fwd, 4),
            "roundtrip_fields_per_s": round(NF / ((inv + fwd) * 1e-3), 1),
            "mean_ms": {"xya_wa": round(inv_mean, 4), "wa_xya": round(fwd_mean, 4)},
            "xya_wa_ms_fresh_result_array": round(fresh, 4),
            "note": f"median of {reps} calls, host arrays reused across calls, PCIe copies included; "
                    "not the bench value"}


SMALL_WORKLOADS = {  # (N, im, jm, levels, gradients)
    "t21l16": (21, 64, 32, 16, False),  # round trip
    "t42l26": (42, 128, 64, 26, True),  # inverse with GradLon/GradLat, forward of a temperature grid
}


def run_small_workload(name: str, steps: int, cpu_seconds: float, seed_cfg: int):
    """The T21L16 round trip and the T42L26 step with gradients: GPU FP64 and
    split-BF16 x6 steps (eager, queue prefilled, events on the stream) and
    the CPU oracle on the host's cores; errors against the FP64 oracle.
    Returns a JSON-ready dict."""
    import torch

    import oracle
    import shtamd
    from shtamd import synth

    n, im, jm, nlev, grads = SMALL_WORKLOADS[name]
    nf = 3 * nlev + 1
    spec_c = synth.state_spectra(n, nlev, cfg=seed_cfg)
    ref = oracle.Plan(n, im, jm)
    out = {"workload": f"T{n}L{nlev}: {nf} fields, {im}x{jm} Gaussian grid"
                       + (", inverse with GradLon/GradLat + forward of a 200-300 K grid" if grads else
                          ", inverse + forward"), "forms": {}}
    # the temperature grid on the oracle's Gauss latitudes (the GPU plan's own
    # agree to ~1e-16); its forward transform is part of the step
    temp = synth.temperature_grid(oracle.gauss(jm)[2], im, nf, cfg=seed_cfg) if grads else None
    for form in ("fp64", "bf16x6"):
        plan = shtamd.Plan(n, im, jm, nf, precision=PRECISIONS[form],
                           flags=shtamd.FLAG_GRAD if grads else 0)
        plan.autotune(nf, reps=10)
        s_in = torch.from_numpy(shtamd.spec_to_real(spec_c)).cuda()
        g = [torch.empty((nf, jm, im), dtype=torch.float64, device="cuda") for _ in range(3)]
        s_out = torch.empty_like(s_in)
        t_in = torch.from_numpy(np.ascontiguousarray(temp)).cuda() if grads else None
        stream = torch.cuda.current_stream()

        def step():
            if grads:
                plan.spec_to_grad(s_in, g[0], g[1], g[2], nf, stream)
                plan.grid_to_spec(t_in, s_out, nf, stream)
            else:
                plan.spec_to_grid(s_in, g[0], nf, stream)
                plan.grid_to_spec(g[0], s_out, nf, stream)

        for _ in range(5):
            step()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda._sleep(SPIN_CYCLES)
        e0.record(stream)
        for _ in range(steps):
            step()
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / steps
        err = {"grid_rel_l2": rel(g[0].cpu().numpy(), ref.spec_to_grid(spec_c))}
        if grads:
            err["gradlon_rel_l2"] = rel(g[1].cpu().numpy(), ref.spec_to_grid(spec_c, kind=1))
            err["gradlat_rel_l2"] = rel(g[2].cpu().numpy(), ref.spec_to_grid(spec_c, kind=2))
            err["forward_rel_l2"] = rel(shtamd.real_to_spec(s_out.cpu().numpy()), ref.grid_to_spec(temp))
        else:
            err["roundtrip_rel_l2"] = rel(shtamd.real_to_spec(s_out.cpu().numpy()), spec_c)
        out["forms"][form] = {"fields_per_s": round(nf / (ms * 1e-3), 1), "ms_per_step": round(ms, 5),
                              "shapes": [SHAPE_NAMES[x] for x in plan.get_shapes(nf)], "err_vs_fp64_oracle": err}
        plan.close()

    def cpu(threads, seconds):
        t0, reps = time.perf_counter(), 0
        while True:
            if grads:
                for kind in range(3):
                    ref.spec_to_grid(spec_c, kind=kind, nthreads=threads)
                ref.grid_to_spec(temp, nthreads=threads)
            else:
                ref.grid_to_spec(ref.spec_to_grid(spec_c, nthreads=threads), nthreads=threads)
            reps += 1
            dt = time.perf_counter() - t0
            if dt >= seconds and reps >= 3:
                return round(reps * nf / dt, 1)

    threads = host_cpu_info()["usable_cpus"]
    out["cpu_oracle"] = {"fields_per_s": cpu(threads, cpu_seconds), "threads": threads,
                         "single_core_fields_per_s": cpu(1, max(1.0, cpu_seconds / 4))}
    return out


def run_t170_workload(nf: int, steps: int, seed_cfg: int = 5):
    """The T170 round trip (512x256 Gaussian grid, the largest truncation of
    the ensemble and serving mixes) at nf fields: FP64 and split-BF16 x6 with
    the built-in shapes (eager, queue prefilled, events on the stream), the
    per-kernel times from the stage timers, and for the split form the
    inverse Legendre step in each of its inverse shapes.  Errors against the
    FP64 oracle on the first 16 fields.  Returns a JSON-ready dict."""
    import torch

    import oracle
    import shtamd
    from shtamd import synth

    n, im, jm = 170, 512, 256
    spec_c = synth.spectra(n, nf, cfg=seed_cfg)
    ref = oracle.Plan(n, im, jm)
    k = min(nf, 16)
    g_ref = ref.spec_to_grid(spec_c[:, :k])
    out = {"metric": "T170 SHT fields/s (inverse + forward, 512x256 Gaussian grid)",
           "workload": f"T170 round trip: {nf} fields, {im}x{jm} Gaussian grid", "fields": nf, "forms": {}}
    s_in = torch.from_numpy(shtamd.spec_to_real(spec_c)).cuda()
    grid = torch.empty((nf, jm, im), dtype=torch.float64, device="cuda")
    s_out = torch.empty_like(s_in)
    stream = torch.cuda.current_stream()

    def timed(fn, reps):
        for _ in range(3):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda._sleep(SPIN_CYCLES)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    for form in ("fp64", "bf16x6"):
        plan = shtamd.Plan(n, im, jm, nf, precision=PRECISIONS[form])

        def step():
            plan.spec_to_grid(s_in, grid, nf, stream)
            plan.grid_to_spec(grid, s_out, nf, stream)

        ms = timed(step, steps)
        plan.set_timing(True)
        timed(step, max(1, steps // 4))
        st = plan.stage_times()
        plan.set_timing(False)
        err = {"grid_rel_l2": rel(grid[:k].cpu().numpy(), g_ref),
               "roundtrip_rel_l2": rel(shtamd.real_to_spec(s_out.cpu().numpy())[:, :k], spec_c[:, :k])}
        f = {"fields_per_s": round(nf / (ms * 1e-3), 1), "ms_per_step": round(ms, 5),
             "kernels_us": {key: round(v[0] / v[1] * 1e3, 2) for key, v in st.items() if v[1]},
             "shapes": [SHAPE_NAMES[x] for x in plan.get_shapes(nf)], "err_vs_fp64_oracle": err}
        if form != "fp64":
            inv = {}
            for name, sh in (("ksplit", shtamd.SHAPE_KSPLIT), ("tile", shtamd.SHAPE_TILE),
                             ("staged", shtamd.SHAPE_STAGED)):
                plan.set_shapes(nf, sh, shtamd.SHAPE_AUTO)
                plan.set_timing(True)
                timed(lambda: plan.spec_to_grid(s_in, grid, nf, stream), max(1, steps // 4))
                v = plan.stage_times()["leg_inv"]
                plan.set_timing(False)
                inv[name] = round(v[0] / v[1] * 1e3, 2)
            f["leg_inv_us_by_shape"] = inv
        out["forms"][form] = f
        plan.close()
    return out


def copy_calibration(nbytes: int = 64 << 20, reps: int = 3):
    """A plain device copy of known size: under rocprofv3 --pmc FETCH_SIZE and
    --pmc WRITE_SIZE, what the counters read per dispatch against the bytes
    actually moved gives the factors that turn them into HBM bytes."""
    import torch

    a = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    return {"workload": "copycal", "bytes_per_copy": nbytes, "copies": reps}


def traffic_from_profile(form: str, kernel_prefix: str):
    """HBM bytes per dispatch of a kernel from the committed rocprofv3 summary
    profiles/t85l32_<form>_hbm_traffic.md (FETCH_SIZE/WRITE_SIZE runs scaled by a
    copy of known size), or None when that summary is absent."""
    path = os.path.join(ROOT, "profiles", f"t85l32_{form}_hbm_traffic.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) >= 7 and cells[0].startswith(kernel_prefix):
            try:
                return int(float(cells[6]) * 1e6)
            except ValueError:
                return None
    return None


def mfma_from_profile(form: str):
    """Measured matrix-core utilisation of the Legendre kernels from the
    committed rocprofv3 counter summary profiles/t85l32_<form>_mfma.md
    (SQ_VALU_MFMA_BUSY_CYCLES over the SIMD-cycles of the profiled duration),
    {stage: {"kernel": name, "util": fraction}}, or None when it is absent."""
    path = os.path.join(ROOT, "profiles", f"t85l32_{form}_mfma.md")
    if not os.path.exists(path):
        return None
    want = {"leg_inv": "leg_inv_" + ("f64" if form == "fp64" else "mixed"),
            "leg_fwd": "leg_fwd_" + ("f64" if form == "fp64" else "mixed")}
    out = {}
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) < 6 or not cells[5].endswith("%"):
            continue
        for stage, prefix in want.items():
            if cells[0].startswith(prefix) and stage not in out:
                out[stage] = {"kernel": cells[0], "util": round(float(cells[5].rstrip(" %")) / 100, 4)}
    return out or None


KERNEL_NAMES = {"fft_inv": "fft_inv_kernel", "fft_fwd": "fft_fwd_kernel"}
SHAPE_NAMES = ("ksplit", "wave", "tile", "tile_split", "staged", "fused")


def legendre_kernel_name(stage: str, fp64: bool, shape: int) -> str:
    """Kernel name (as rocprof prints it, up to the template list) of a
    Legendre stage in a given shape."""
    if not fp64 and stage == "leg_inv" and shape == 5:
        return "inv_fused_kernel"  # Legendre step and FFT in one launch
    if fp64:
        suffix = ("_f64_kernel", "_f64_wave_kernel", "_f64_tile_kernel", "_f64_tile_kernel",
                  "_f64_staged_kernel")[shape]
    else:
        suffix = ("_mixed_kernel", "_mixed_wave_kernel", "_mixed_kernel")[shape]  # TILE: 2 column tiles
    return stage + suffix

ENSEMBLE_GRIDS = {42: (128, 64), 85: (256, 128), 170: (512, 256)}


def run_ensemble(precision: str, nmember: int, steps: int, warmup: int, seed: int = 20261202,
                 shapes_only: bool = False):
    """Ensemble of nmember members, truncation uniform in {T42, T85, T170},
    levels uniform in {16, 32, 48} (3L+1 fields each), one graph launch per
    round trip; returns a JSON-ready dict.  shapes_only: the autotuner picks
    one Legendre shape per direction and group, with the built-in wavenumber
    classes; otherwise FP64 groups also tune the split shape's classes."""
    import torch

    import oracle
    import shtamd
    from shtamd import hipmem, synth

    rng = np.random.default_rng(seed)
    Ts = rng.choice([42, 85, 170], size=nmember)
    Ls = rng.choice([16, 32, 48], size=nmember)
    members = [(int(T), *ENSEMBLE_GRIDS[int(T)], 3 * int(L) + 1) for T, L in zip(Ts, Ls)]
    b = shtamd.Batch(members, precision=PRECISIONS[precision])
    specs = []
    for j, (N, im, jm, nf) in enumerate(members):
        s = synth.state_spectra(N, (nf - 1) // 3, cfg=3, member=j)
        spec_in, _, _, ld = b.member_ptrs(j)
        hipmem.put_member_spectra(spec_in, ld, s)
        specs.append(s)
    # per-group Legendre shapes from the autotuner, then the graph is captured
    b.autotune(reps=5, shapes_only=shapes_only)
    stream = torch.cuda.current_stream()
    for _ in range(max(1, warmup)):
        b.run(shtamd.DIR_ROUNDTRIP, stream)
    torch.cuda.synchronize()
    times = {}
    for mode in ("graph", "eager"):
        fn = b.run if mode == "graph" else b.run_eager
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda._sleep(SPIN_CYCLES)
        e0.record(stream)
        for _ in range(steps):
            fn(shtamd.DIR_ROUNDTRIP, stream)
        e1.record(stream)
        torch.cuda.synchronize()
        times[mode] = e0.elapsed_time(e1) / steps
    # per-member error against the FP64 oracle
    errs = []
    plans = {}
    for j, (N, im, jm, nf) in enumerate(members):
        spec_in, grid, spec_out, ld = b.member_ptrs(j)
        K = (N + 1) * (N + 2) // 2
        g = hipmem.from_device(grid, (nf, jm, im))
        s_rt = hipmem.get_member_spectra(spec_out, ld, K, nf)
        ref = plans.setdefault(N, oracle.Plan(N, im, jm))
        errs.append(max(rel(g, ref.spec_to_grid(specs[j])), rel(s_rt, specs[j])))
    b.close()
    nf_total = sum(m[3] for m in members)
    best = min(times.values())
    return {
        "metric": "ensemble SHT fields/s (64 members, T42/T85/T170 x L16/32/48, one graph launch)",
        "value": round(nf_total / (best * 1e-3), 1), "unit": "fields/s", "precision": precision,
        "tuning": "shapes per direction" if shapes_only else "shapes and wavenumber classes",
        "members": nmember, "fields": nf_total,
        "members_per_truncation": {f"T{t}": int((Ts == t).sum()) for t in (42, 85, 170)},
        "ms_per_round_trip": {k: round(v, 4) for k, v in times.items()},
        "max_member_rel_l2_vs_fp64_oracle": max(errs),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--precision", default="auto", choices=sorted(PRECISIONS) + ["auto"],
                    help="auto: run every form; the headline is split-BF16 (six products, <=1e-6)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--workload", default="t85l32", choices=["t85l32", "ensemble", "serve", "copycal", "small",
                                                              "t170"],
                    help="ensemble: the 64-member mixed-truncation batch; serve: request streams at "
                         "100-1000 req/s (separate JSON lines); copycal: a plain 64 MiB device copy, "
                         "three dispatches, the calibration run for rocprofv3's FETCH_SIZE/WRITE_SIZE; "
                         "small: the T21L16 round trip and the T42L26 step with gradients, GPU and CPU oracle; "
                         "t170: the T170 round trip at --fields fields (one JSON line each)")
    ap.add_argument("--fields", default="145,776",
                    help="field counts of the t170 workload (comma-separated)")
    ap.add_argument("--serve-requests", type=int, default=500,
                    help="requests in the 500 req/s serving run of the headline line (0 = skip)")
    ap.add_argument("--members", type=int, default=64)
    ap.add_argument("--shapes", default="auto", choices=["auto", "default", "fused"],
                    help="Legendre kernel shapes: autotuned per form (auto) or the built-in choice "
                         "(default; what the rocprof summaries use, without the tuner's trial launches); "
                         "fused: the built-in choice with the fused inverse for the mixed forms")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "the transform runs on one GPU; --gpus above 1 is not supported "
                                   "(the transform would shard by field batch, one slab of levels per GPU)"}))
        sys.exit(2)

    if args.workload == "copycal":
        print(json.dumps(copy_calibration()))
        return
    if args.workload == "t170":
        for nf in [int(x) for x in args.fields.split(",")]:
            print(json.dumps(run_t170_workload(nf, max(1, args.steps // 10))), flush=True)
        return
    if args.workload == "small":
        for k, name in enumerate(SMALL_WORKLOADS):
            print(json.dumps(run_small_workload(name, args.steps, min(args.cpu_seconds, 5.0), k)), flush=True)
        return

    import oracle

    if args.workload == "ensemble":
        # the headline form (split-BF16 x6), then FP64 tuned two ways: one
        # shape per direction, and shapes with the split shape's wavenumber
        # classes (FP64 only); one JSON line each
        precs = ["bf16x6", "fp64"] if args.precision == "auto" else [args.precision]
        for prec in precs:
            for shapes_only in ((False,) if prec != "fp64" else (True, False)):
                print(json.dumps(run_ensemble(prec, args.members, max(1, args.steps // 10), args.warmup,
                                              shapes_only=shapes_only)), flush=True)
        return
    if args.workload == "serve":
        prec = "bf16x6" if args.precision == "auto" else args.precision
        # 100-1000 req/s: the config's range (requests mostly run alone); 5000
        # and 20000 req/s: arrivals overlap and the server batches them
        res = run_serve(prec, [100, 250, 500, 1000, 5000, 20000], lambda r: max(200, min(2000, 2 * r)))
        res["metric"] = "SHT serving p50/p99 latency (ms) under Poisson load, T21/T42/T85/T170 mix"
        print(json.dumps(res))
        return

    forms = list(PRECISIONS) if args.precision == "auto" else sorted({"fp64", args.precision})
    headline = "bf16x6" if args.precision == "auto" else args.precision
    ref = oracle.Plan(N, IM, JM)
    results = {}
    for form in forms:
        res = run_gpu(PRECISIONS[form], args.steps, args.warmup, autotune=args.shapes == "auto",
                      fused=args.shapes == "fused")
        g_ref = ref.spec_to_grid(res["spec_in"])
        res["err"] = {"grid_rel_l2": rel(res["grid"], g_ref),
                      "spec_rel_l2": rel(res["spec"], ref.grid_to_spec(res["grid"])),
                      "roundtrip_rel_l2": rel(res["spec"], res["spec_in"])}
        res["best_ms"] = min(x for x in (res["eager_ms"], res["graph_ms"]) if x is not None)
        results[form] = res

    res = results[headline]
    prec = PRECISIONS[headline]
    value = NF / (res["best_ms"] * 1e-3)

    # roofline of the dominant kernel of the headline form
    tb = table_bytes(prec)
    alg = algorithmic(tb, NF, 16 if prec == 0 else 8)
    hshapes = results[headline]["shapes"]
    if prec != 0 and hshapes[0] == 5:
        # the fused inverse: one launch, timed as the Legendre stage, doing the
        # work of both inverse stages (the intermediate is still written and
        # read once), so its algorithmic bytes are their sum
        (lb, lf), (fb, ff) = alg["leg_inv"], alg["fft_inv"]
        alg["leg_inv"] = (lb + fb, lf + ff)
    kms = {k: v for k, v in res["kernel_ms"].items() if v}
    dom = max(kms, key=kms.get)
    bytes_, flops = alg[dom]
    achieved = bytes_ / (kms[dom] * 1e-3) / 1e9
    if dom in KERNEL_NAMES:
        # the forward FFT carries a third template argument (direct first pass)
        kprefix = f"{KERNEL_NAMES[dom]}<{IM},{'double' if prec == 0 else 'float'}"
        kname = kprefix + ">"
        traffic = traffic_from_profile(headline, kprefix)
    else:
        kname = legendre_kernel_name(dom, prec == 0, hshapes[0 if dom == "leg_inv" else 1])
        traffic = traffic_from_profile(headline, kname + "<")
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": kname,
                "kernel_us": round(kms[dom] * 1e3, 2), "algorithmic_bytes": int(bytes_)}
    per_form = {}
    for form, r in results.items():
        t = {"fields_per_s": round(NF / (r["best_ms"] * 1e-3), 1), "ms_per_step": round(r["best_ms"], 5),
             "eager_ms": round(r["eager_ms"], 5),
             "graph_ms": None if r["graph_ms"] is None else round(r["graph_ms"], 5),
             "err_vs_fp64_oracle": r["err"], "kernels_us": {},
             "legendre_shapes": {"leg_inv": SHAPE_NAMES[r["shapes"][0]], "leg_fwd": SHAPE_NAMES[r["shapes"][1]]}}
        a = algorithmic(table_bytes(PRECISIONS[form]), NF, 16 if form == "fp64" else 8)
        for k, ms in r["kernel_ms"].items():
            if not ms:
                continue
            b, f = a[k]
            t["kernels_us"][k] = round(ms * 1e3, 2)
            if k.startswith("leg"):
                peak = FP64_PEAK_TFLOPS if form == "fp64" else BF16_PEAK_TFLOPS
                t.setdefault("legendre_mfma_frac", {})[k] = round(f / (ms * 1e-3) / 1e12 / peak, 4)
        # measured by rocprofv3 counters (matrix-core busy cycles), built-in shapes
        t["legendre_mfma_util_measured"] = mfma_from_profile(form)
        per_form[form] = t
    fp64_v = per_form["fp64"]["fields_per_s"]

    host_res = run_host_dropin(prec)

    serve_res = None
    if args.serve_requests > 0:
        serve_res = run_serve(headline, [500], lambda r: args.serve_requests)["rates"]["500"]

    cpu = host_cpu_info()
    threads = cpu["usable_cpus"]  # every core this process may run on
    cpu_v, cpu_n, cpu_dt = cpu_baseline(res["spec_in"], args.cpu_seconds, threads)
    cpu1_v, cpu1_n, cpu1_dt = cpu_baseline(res["spec_in"], max(2.0, args.cpu_seconds / 4), 1)

    line = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "fields/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(res["best_ms"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {0: "f64", 1: "bf16x2pieces/3products/f32-slice/f64-sum",
                  2: "f16x2pieces-scaled/3products/f32-slice/f64-sum",
                  3: "bf16x3pieces/6products/f32-slice/f64-sum"}[prec],
        "data": "synthetic (seeded CN(0,n^-3) spectra, master seed 20261202)",
        "config": {"workload": "T85L32 forward+inverse SHT: 97 fields, 256x128 Gaussian grid, N=85",
                   "fields_per_step": NF, "precision": headline,
                   "timing": "graph" if res["graph_ms"] is not None and res["graph_ms"] <= res["eager_ms"] else "eager"},
        "err_vs_fp64_oracle": res["err"],
        "speedup_vs_fp64_form": round(value / fp64_v, 3),
        "forms": per_form,
        "serve_500rps": serve_res,
        "host_dropin": host_res,
        "roofline": roofline,
        "cpu_baseline": {"value": round(cpu_v, 1), "unit": "fields/s", "cores": threads, "kind": "port",
                         "sample": f"oracle round trips of the same 97-field T85L32 batch, {cpu_n} reps "
                                   f"in {cpu_dt:.1f} s, {threads} OpenMP threads (one per usable CPU)",
                         "single_core_value": round(cpu1_v, 1), "host": cpu},
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
