"""Benchmark: volumes/s of the RF -> volume path on one MI355X.

Workload (default): 128-element row-column array, 128 Hadamard-encoded
transmits x 128 channels x 4096 int16 samples (128 MiB raw frame, already in
HBM), channel remap + int8-MFMA decode, 64-tap match filter + Hilbert, HERCULES
3-D DAS to 128^3 complex64 with linear interpolation, F# 1.5 Hann.  One step
is one full pass raw -> volume.  Prints one JSON line.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0
FLOP_PER_TAP = {"linear": 34, "cubic": 65}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def _cpu_model():
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(acq, interp, runs=3):
    """FP64 oracle on the box's cores: the full frame's convert + decode and
    Hilbert, and the DAS of 2 of the 128 depth planes (iz = 48, 80) scaled to
    the volume; one warm-up pass, then the median of `runs` timed passes per
    stage."""
    import numpy as np

    import oracle
    from cuda_ultrasound_toolkit_amd import _lib as L
    oracle.load()
    bp = acq.bp
    S, Cc, T = bp.dec_data_dim[0], bp.dec_data_dim[1], bp.dec_data_dim[2]
    nz = bp.output_points[2]
    planes = [48, 80]
    cmap = np.array(bp.channel_mapping[:Cc], np.int16)
    H = oracle.hadamard(T, bp.decode)
    h = np.array(bp.rf_filter[: bp.filter_length], np.float64)

    def one(n_planes):
        t0 = time.perf_counter()
        dec = oracle.decode(oracle.convert(acq.raw, L.TYPE_I16, cmap, bp.rf_raw_dim[0], T, Cc, S), H)
        t1 = time.perf_counter()
        iq = oracle.hilbert(dec.reshape(T * Cc, S), h).reshape(T, Cc, S)
        t2 = time.perf_counter()
        oracle.das(iq, bp, cubic=(interp == "cubic"), planes=planes[:n_planes])
        t3 = time.perf_counter()
        return t1 - t0, t2 - t1, (t3 - t2) * nz / n_planes

    one(1)  # warm-up (page-in, thread pool start)
    t = np.array([one(len(planes)) for _ in range(runs)])
    med = np.median(t, axis=0)
    total = float(med.sum())
    try:
        threads = len(os.sched_getaffinity(0))
    except AttributeError:
        threads = os.cpu_count() or 1
    threads = int(os.environ.get("OMP_NUM_THREADS", threads))
    return {"value": 1.0 / total, "unit": "volumes/s", "cores": threads, "kind": "port",
            "cpu": _cpu_model(), "threads": threads, "runs": runs,
            "sample": (f"FP64 oracle (OpenMP, {threads} threads): convert + decode and Hilbert of the full frame, "
                       f"DAS of {len(planes)}/{nz} depth planes scaled to the volume; median of {runs} runs after a "
                       f"warm-up; stage s/volume decode {med[0]:.2f}, hilbert {med[1]:.2f}, das {med[2]:.1f}")}


MOTION_SEEDS = (20261203, 777, 4242, 99)


def motion_leg(dev, stream, seeds=MOTION_SEEDS, repeats=5):
    """3-D block matching of 128^3 envelope pairs (rigid shift U[-6, 6] voxels per
    axis, 2 % decorrelation; one pair per seed): full +-8 ZNCC search vs the
    learned proposal + +-2 refine with full-search fallback, at three peak
    thresholds.  Times are medians per pair; the figures are means over the
    pairs (agreement also as the minimum).  Besides agreement with the full
    search, both are scored against the known rigid shift: the fraction of
    blocks with full-search peak >= 0.6 whose displacement is within one voxel
    of it on every axis."""
    import numpy as np
    import torch

    from cuda_ultrasound_toolkit_amd import api, motion_data
    from cuda_ultrasound_toolkit_amd import motion_learned as ml

    assert api.set_das_options(cubic=False)
    params, meta = ml.load_params()
    assert api.pnet_set_input_mode(meta.get("input_mode", ml.INPUT_MODE))
    dp = torch.from_numpy(params).to(dev)
    thresholds = sorted({0.6, 0.65, float(meta.get("threshold", 0.6))})

    def best_ms(fn):
        fn()
        torch.cuda.synchronize()
        t = []
        for _ in range(repeats):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record(stream)
            fn()
            b.record(stream)
            torch.cuda.synchronize()
            t.append(a.elapsed_time(b))
        return float(np.median(t))

    full_ms, confident, shifts, full_truth = [], [], [], []
    per_thr = {thr: [] for thr in thresholds}
    nb = 0
    for seed in seeds:
        log(f"bench: motion pair seed {seed}")
        shift = motion_data.random_shifts(np.random.default_rng(seed), 1)[0]
        shifts.append(np.round(shift, 2).tolist())
        tpl, src, _ = motion_data.envelope_pair(seed, shift, dev)
        nz, ny, nx = tpl.shape
        nb = ((nx - 16) // 8 + 1) * ((ny - 16) // 8 + 1) * ((nz - 16) // 8 + 1)
        fd = torch.zeros((nb, 3), dtype=torch.int32, device=dev)
        fp = torch.zeros(nb, dtype=torch.float32, device=dev)
        ld = torch.zeros_like(fd)
        lp = torch.zeros_like(fp)
        ms_full = best_ms(lambda: api.ncc3d_search(tpl, src, 8, fd, fp))
        full = fd.cpu().numpy()
        full_ms.append(ms_full)
        conf = fp.cpu().numpy() >= 0.6
        confident.append(float(np.mean(conf)))

        def truth(d):  # confident blocks whose displacement is within 1 voxel (every axis) of the rigid shift
            return float(np.mean(np.abs(d[conf] - np.asarray(shift)[None, :]).max(axis=1) < 1.0))

        full_truth.append(truth(full))
        for thr in thresholds:
            _, nf = api.motion3d_learned(dp, tpl, src, 8, ld, lp, threshold=thr, count_fallbacks=True)
            got = ld.cpu().numpy()
            agree = float(np.mean(np.all(got == full, axis=1)))
            ms = best_ms(lambda: api.motion3d_learned(dp, tpl, src, 8, ld, lp, threshold=thr))
            per_thr[thr].append((ms, ms_full / ms, agree, nf / nb, truth(got)))
    mf = float(np.mean(full_ms))
    out = {"workload": f"{len(seeds)} pairs of 128^3 envelope volumes (shifts {shifts} voxels), {nb} blocks of 16^3 "
                       "at stride 8 per pair, +-8 search; means over the pairs",
           "full_ms": round(mf, 4), "full_blocks_per_s": round(nb / (mf * 1e-3)),
           "full_tflops": round(nb * 2 * 4096 * 17 ** 3 / (mf * 1e-3) / 1e12, 2),
           "confident_fraction": round(float(np.mean(confident)), 4),
           "full_within1_of_shift": round(float(np.mean(full_truth)), 4)}
    for thr in thresholds:
        a = np.array(per_thr[thr])
        out[f"learned_thr{thr:.2f}"] = {"ms": round(float(a[:, 0].mean()), 4),
                                        "blocks_per_s": round(nb / (float(a[:, 0].mean()) * 1e-3)),
                                        "speedup_vs_full": round(float(a[:, 1].mean()), 3),
                                        "agreement": round(float(a[:, 2].mean()), 4),
                                        "agreement_min": round(float(a[:, 2].min()), 4),
                                        "fallback_rate": round(float(a[:, 3].mean()), 4),
                                        "within1_of_shift": round(float(a[:, 4].mean()), 4)}
    log(f"bench: motion {json.dumps(out)}")
    return out


def sequences_leg(dev, stream, repeats=3):
    """FORCES and UFORCES (2-D, 32-element linear array) and RCA_TPW (3-D, 32
    tilted plane waves x 128 columns -> 128^3) DAS kernels on device-resident
    IQ, and the bandpass input stage at the full frame size."""
    import ctypes as C
    import torch

    from cuda_ultrasound_toolkit_amd import _lib as L, api, configs
    lib = L.load()
    out = {}
    assert api.set_das_options(cubic=True)  # the sequences' interpolation, as labelled below
    cases = {
        "forces_2d": (configs.linear_2d(seed=1, generate=False),
                      "FORCES: 32 tx x 32 ch x 1024 samples -> 128 x 256, cubic, no apodization"),
        "uforces_2d": (configs.uforces_2d(seed=6, T=32, S=1024, generate=False),
                       "UFORCES: flash + 31 sparse transmits x 32 ch x 1024 samples -> 128 x 256, cubic"),
        "tpw_3d": (configs.tpw_3d(seed=2, T=32, C=128, S=4096, n=128, generate=False, z_range=(5e-3, 60e-3),
                                  half_width=19e-3, box_half=19e-3),
                   "RCA_TPW: 32 angles x 128 ch x 4096 samples -> 128^3, cubic, F# 1.5 Hann"),
    }
    for name, (acq, desc) in cases.items():
        bp = acq.bp
        S, Cc, T = bp.dec_data_dim[0], bp.dec_data_dim[1], bp.dec_data_dim[2]
        n = bp.output_points
        g = torch.Generator(device=dev).manual_seed(11)
        d_iq = torch.randn(T * Cc * S * 2, device=dev, generator=g)
        d_vol = torch.zeros(n[0] * n[1] * n[2] * 2, dtype=torch.float32, device=dev)
        assert api.das_device(d_iq, d_vol, bp)
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for _ in range(repeats):
            assert api.das_device(d_iq, d_vol, bp)
        b.record(stream)
        torch.cuda.synchronize()
        ms = a.elapsed_time(b) / repeats
        taps = C.c_ulonglong()
        assert lib.cut_das_active_taps(C.byref(bp), C.byref(taps))
        out[name] = {"workload": desc, "ms": round(ms, 4), "active_taps": taps.value,
                     "gtaps_per_s": round(taps.value / (ms * 1e-3) / 1e9, 2)}
        del d_iq, d_vol
    # bandpass input stage at the full frame size: (I, Q) int16 pairs -> demodulated complex, 64-tap FIR
    import numpy as np
    T, Cc, S, taps_n = 128, 128, 4096, 64
    g = torch.Generator(device=dev).manual_seed(12)
    d_raw = (torch.randn(Cc * 2 * T * S, device=dev, generator=g) * 3000).to(torch.int16)
    d_out = torch.empty(T * Cc * S * 2, dtype=torch.float32, device=dev)
    assert api.rf_init((2 * T * S, Cc), (S, Cc, T), L.NONE)
    assert api.rf_set_channel_mapping(np.arange(Cc, dtype=np.int16))
    assert api.rf_set_match_filter(np.hanning(taps_n + 2)[1:-1].astype(np.float32))
    run = lambda: api.rf_convert_demod(d_raw, L.TYPE_I16, d_out, 5e6, 20e6)
    assert run()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(stream)
    for _ in range(repeats):
        assert run()
    b.record(stream)
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / repeats
    nbytes = (4 + 8) * T * Cc * S
    out["bandpass_demod"] = {"workload": "bandpass (I, Q) int16 pairs 128 tx x 128 ch x 4096 -> demodulated complex, "
                                         "64-tap centred FIR", "ms": round(ms, 4),
                             "GBs": round(nbytes / (ms * 1e-3) / 1e9, 1),
                             "frac_hbm": round(nbytes / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4)}
    assert api.rf_set_match_filter(None)
    del d_raw, d_out
    log(f"bench: sequences {json.dumps(out)}")
    return out


def batch_leg(dev, n, seed=4, repeats=3):
    """n frames of uneven size (transmits {16..128}, 128 channels, samples
    U{1024..4096}, cube side U{64..160}): one graph-captured batch with a
    persistent DAS launch vs the same frames one after another."""
    import numpy as np
    import torch

    from cuda_ultrasound_toolkit_amd import api, configs

    log(f"bench: synthesizing {n} uneven frames")
    t0 = time.perf_counter()
    acqs = configs.uneven_frames(seed=seed, n=n)
    log(f"bench: synthesis {time.perf_counter() - t0:.1f} s")
    assert api.set_das_options(cubic=False)
    raws = [torch.from_numpy(a.raw).to(dev) for a in acqs]
    vols = [torch.zeros(int(np.prod(a.bp.output_points[:3])) * 2, dtype=torch.float32, device=dev) for a in acqs]
    frames = [(r, v, a.bp) for r, v, a in zip(raws, vols, acqs)]
    res = {}
    for name, mode in (("graph_persistent", 0), ("graph_in_order", 2), ("sequential", 1)):
        ok, _ = api.beamform_batch(frames, mode=mode)  # set-up, tuning, capture
        assert ok
        lat, wall = [], []
        for _ in range(repeats):
            torch.cuda.synchronize()
            t = time.perf_counter()
            ok, ms = api.beamform_batch(frames, mode=mode, latencies=True)
            wall.append(time.perf_counter() - t)
            assert ok
            lat.append(ms)
        gpu_ms = [max(m) for m in lat]
        best = int(np.argmin(gpu_ms))
        m = np.array(lat[best])
        res[name] = {"batch_gpu_ms": round(gpu_ms[best], 3), "volumes_per_s": round(n / (gpu_ms[best] * 1e-3), 2),
                     "wall_ms": round(1000 * min(wall), 3),
                     "latency_p50_ms": round(float(np.percentile(m, 50)), 3),
                     "latency_p99_ms": round(float(np.percentile(m, 99)), 3)}
    ok, _ = api.beamform_batch(frames, mode=0)  # the launch shape of the throughput mode
    g, o = api.batch_launch_shape()
    res["launch_shape"] = {"workgroups_per_cu": g, "order": ["heaviest first", "frame by frame"][o]}
    res["speedup"] = round(res["sequential"]["batch_gpu_ms"] / res["graph_persistent"]["batch_gpu_ms"], 3)
    res["workload"] = (f"{n} frames, seed {seed}: transmits {{16,32,64,128}}, 128 channels, samples U{{1024..4096}}, "
                       "50-500 scatterers, cube side U{64..160}, linear interpolation")
    res["total_voxels"] = int(sum(int(np.prod(a.bp.output_points[:3])) for a in acqs))
    assert api.set_das_options(cubic=True)
    log(f"bench: batch {json.dumps(res)}")
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--interp", choices=["linear", "cubic"], default="linear")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--stage-iters", type=int, default=10)
    ap.add_argument("--no-motion", action="store_true", help="skip the 3-D motion leg (learned vs full search)")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-frame uneven batch leg")
    ap.add_argument("--batch-frames", type=int, default=64)
    ap.add_argument("--no-sequences", action="store_true", help="skip the FORCES / UFORCES / RCA_TPW kernel timings")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run one device-to-device copy of the 512 MiB IQ buffer (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); frames shard with no exchange, "
              "see the sharding notes, but no multi-GPU figure is produced.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from cuda_ultrasound_toolkit_amd import _lib as L, api, configs

    lib = L.load()
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    log("bench: synthesizing 128 tx x 128 ch x 4096 frame")
    acq = configs.rca_3d(seed=2)
    bp = acq.bp
    S, Cc, T = bp.dec_data_dim[0], bp.dec_data_dim[1], bp.dec_data_dim[2]
    nx, ny, nz = bp.output_points[0], bp.output_points[1], bp.output_points[2]
    d_raw = torch.from_numpy(acq.raw).to(dev)
    d_vol = torch.zeros(nx * ny * nz * 2, dtype=torch.float32, device=dev)
    # a dedicated (non-null) stream: the library enqueues on it and the timing
    # events are recorded on it
    stream = torch.cuda.Stream(device=dev)
    torch.cuda.set_stream(stream)
    api.set_stream(stream.cuda_stream)
    assert api.set_das_options(cubic=(args.interp == "cubic"))

    log("bench: warmup")
    for _ in range(args.warmup):
        assert api.beamform_device(d_raw, d_vol, bp)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t_wall = time.perf_counter()
    e0.record(stream)
    for _ in range(args.steps):
        assert api.beamform_device(d_raw, d_vol, bp)
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_wall
    ms = e0.elapsed_time(e1) / args.steps
    log(f"bench: {ms:.3f} ms/volume (wall {1000 * wall / args.steps:.3f})")

    # ---- API-level figure: cuda_toolkit::beamform (the reference's entry point)
    # from host buffers: H2D of the raw frame, the pipeline, D2H of the volume,
    # synchronous like the reference; pinned and pageable host memory
    api_level = {}
    fn = getattr(lib, L.SYM_BEAMFORM)
    for kind in ("pinned", "pageable"):
        if kind == "pinned":
            h_raw = torch.from_numpy(acq.raw).pin_memory()
            h_vol = torch.empty(nx * ny * nz * 2, dtype=torch.float32).pin_memory()
            sp_in, sp_out = L.Span(h_raw.data_ptr(), acq.raw.nbytes), L.Span(h_vol.data_ptr(), nx * ny * nz * 8)
        else:
            h_vol_np = np.empty(nx * ny * nz * 2, np.float32)
            sp_in = L.Span(acq.raw.ctypes.data, acq.raw.nbytes)
            sp_out = L.Span(h_vol_np.ctypes.data, h_vol_np.nbytes)
        for _ in range(2):
            assert fn(sp_in, sp_out, C.byref(bp))
        reps = max(5, args.steps // 2)
        t0 = time.perf_counter()
        for _ in range(reps):
            assert fn(sp_in, sp_out, C.byref(bp))
        am = 1000.0 * (time.perf_counter() - t0) / reps
        api_level[kind] = {"ms_per_volume": round(am, 4), "volumes_per_s": round(1000.0 / am, 3)}
    api_level["workload"] = ("cuda_toolkit::beamform on host spans: 128 MiB int16 frame H2D, the same pipeline, "
                             "16 MiB volume D2H, synchronous per call")
    log(f"bench: API level {json.dumps(api_level)}")

    # ---- precision variant: IQ staged as FP16 / BF16 between Hilbert and DAS,
    # coherence factor applied (FP32 accumulation throughout)
    precision = {}
    for name, fmt in (("f32", 0), ("fp16", 1), ("bf16", 2)):
        assert api.set_das_options(cubic=(args.interp == "cubic"), cf_hercules=True, iq_format=fmt)
        assert api.beamform_device(d_raw, d_vol, bp)
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for _ in range(max(3, args.steps // 4)):
            assert api.beamform_device(d_raw, d_vol, bp)
        b.record(stream)
        torch.cuda.synchronize()
        pm = a.elapsed_time(b) / max(3, args.steps // 4)
        precision[name] = {"ms_per_volume": round(pm, 4), "volumes_per_s": round(1000.0 / pm, 3)}
    assert api.set_das_options(cubic=(args.interp == "cubic"), cf_hercules=False, iq_format=0)
    log(f"bench: precision variants {json.dumps(precision)}")

    # ---- per-stage timing on the same data (stage entry points, same stream)
    def timed(fn, iters):
        fn()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for _ in range(iters):
            fn()
        b.record(stream)
        torch.cuda.synchronize()
        return a.elapsed_time(b) / iters

    assert api.rf_init((bp.rf_raw_dim[0], bp.rf_raw_dim[1]), (S, Cc, T), bp.decode)
    assert api.rf_set_channel_mapping(np.array(bp.channel_mapping[:Cc], np.int16))
    assert api.rf_set_match_filter(np.array(bp.rf_filter[: bp.filter_length], np.float32))
    d_dec = torch.zeros(T * Cc * S, dtype=torch.float32, device=dev)
    d_iq = torch.zeros(T * Cc * S * 2, dtype=torch.float32, device=dev)
    it = args.stage_iters
    ms_dec = timed(lambda: api.rf_convert_decode(d_raw, L.TYPE_I16, d_dec, 1, True), it)
    ms_hil = timed(lambda: api.rf_hilbert(d_dec, 1, d_iq), it)
    ms_das = timed(lambda: api.das_device(d_iq, d_vol, bp), max(3, it // 2))
    api.set_decode_mode(1)
    ms_dec_f32 = timed(lambda: api.rf_convert_decode(d_raw, L.TYPE_I16, d_dec, 1, True), it)
    api.set_decode_mode(0)
    other_interp = "cubic" if args.interp == "linear" else "linear"
    api.set_das_options(cubic=(other_interp == "cubic"))
    ms_das_other = timed(lambda: api.das_device(d_iq, d_vol, bp), 3)
    api.set_das_options(cubic=(args.interp == "cubic"))
    if args.copy_calibration:
        d_copy = torch.empty_like(d_iq)
        d_copy.copy_(d_iq)
        torch.cuda.synchronize()
        log(f"bench: calibration copy of {d_iq.numel() * 4} bytes")
    motion = None if args.no_motion else motion_leg(dev, stream)
    sequences = None if args.no_sequences else sequences_leg(dev, stream)
    batch = None if args.no_batch else batch_leg(dev, args.batch_frames)
    api.set_stream(None)

    taps = C.c_ulonglong()
    assert lib.cut_das_active_taps(C.byref(bp), C.byref(taps))
    tcs = T * Cc * S
    das_tflops = taps.value * FLOP_PER_TAP[args.interp] / (ms_das * 1e-3) / 1e12
    das_tflops_other = taps.value * FLOP_PER_TAP[other_interp] / (ms_das_other * 1e-3) / 1e12
    dec_gbs = (2 + 4) * tcs / (ms_dec * 1e-3) / 1e9
    hil_gbs = (4 + 8) * tcs / (ms_hil * 1e-3) / 1e9

    traffic = None
    tfile = ROOT / "profiles" / "traffic.json"
    if tfile.exists():
        traffic = json.loads(tfile.read_text()).get(f"das_{args.interp}")

    cpu = None
    if not args.no_cpu_baseline:
        log("bench: cpu baseline (FP64 oracle sample)")
        cpu = cpu_baseline(acq, args.interp)

    value = 1000.0 / ms
    out = {
        "metric": "volumes/s at 128ch x 128tx -> 128^3 (raw int16 in HBM -> complex64 volume)",
        "value": round(value, 3),
        "unit": "volumes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded point-scatterer RCA frame, Hadamard-128 encoded)",
        "config": {"workload": "rca_3d: 128 tx x 128 ch x 4096 int16 -> decode -> 64-tap filter + Hilbert "
                               f"-> HERCULES DAS 128^3, {args.interp} interpolation, F# 1.5 Hann",
                   "interp": args.interp, "decode": "int8 MFMA hi/lo (bit-exact)"},
        "roofline": {"bound": "valu", "kernel": "hercules_kernel + hercules_reduce (DAS launch)",
                     "achieved": round(das_tflops, 3), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(das_tflops / PEAK_FP32_TFLOPS, 4), "traffic": traffic,
                     "work": f"{taps.value} active taps x {FLOP_PER_TAP[args.interp]} FLOP",
                     "kernel_ms": round(ms_das, 4)},
        "stages_ms": {"decode_i8": round(ms_dec, 4), "decode_f32": round(ms_dec_f32, 4),
                      "hilbert": round(ms_hil, 4), f"das_{args.interp}": round(ms_das, 4),
                      f"das_{other_interp}": round(ms_das_other, 4)},
        "stages_roofline": {"decode_i8_GBs": round(dec_gbs, 1), "decode_i8_frac_hbm": round(dec_gbs / PEAK_HBM_GBS, 4),
                            "hilbert_GBs": round(hil_gbs, 1), "hilbert_frac_hbm": round(hil_gbs / PEAK_HBM_GBS, 4),
                            f"das_{other_interp}_TFLOPs": round(das_tflops_other, 3)},
        "cpu_baseline": cpu,
        "api_level": api_level,
        "precision_cf_on": precision,
        "motion": motion,
        "batch": batch,
        "das_sequences": sequences,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
