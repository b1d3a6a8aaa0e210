#!/usr/bin/env python3
"""Benchmark of the RSF level-set evolution on one MI355X.

Headline workload (the plain drop-in configuration): a 512^3 uint8 tube
phantom (300 tubes, radius 2-6, +-20 x ramp, noise sigma 15, seed 2), fp32 phi0
= signed distance to the thresholded sigma-1 blur, sigma = 3 (21 taps),
lambda1 = lambda2 = 1, nu = 65.025, mu = 1, dt = 0.1, epsilon = 1, 500
iterations.  A step is one RSF iteration over the whole volume; `value` is
voxel-iterations per second with image and phi resident in HBM, in the literal
mode (the reference's four blurred channels and operation order; bit-identical
to the oracle).

Also reported on the same line:
  fast_mode     the same run in the fast mode (two blurred channels, H = 1/2 + G;
                meets the one-step gate, not bit-identical);
  roofline      the dominant kernel's algorithmic bytes per launch over its
                mean launch time (HIP events on the launch stream), against
                MI355X's 8 TB/s; `traffic` = rocprofv3-measured bytes/launch
                from profiles/ when present;
                `valu` = sweep 1's minimum FP32 FMA count (the separable blurs)
                over its launch time, against the 78.6 TFLOP/s VALU peak;
  loop_roofline 21 B per voxel-iteration over the whole timed loop;
  batched       64 uneven blocks (sides from {96,160,192,256}, seed 1000+b),
                300 iterations: one persistent launch per sweep per
                iteration in a graph, vs the same blocks one after another;
  adjoint       the differentiated batch: 8 x 128^3 (seeds 4000+b), 32 unrolled
                steps with checkpoints every 8, BCE loss; forward and backward;
  large         1024x1024x768 (2000 tubes, radius 1.5-8, noise 20, seed 5) for
                --steps iterations with the default kernels and with the
                autotuned ones;
  cpu_baseline  the oracle (CPU restatement of the reference pipeline,
                unfused, FP32) on the same 512^3 volume for a bounded number
                of iterations on this host's cores, one thread on a 256^3 crop,
                and the 128^3 CPU configuration (40 tubes, seed 1) run in full
                for 100 iterations; `host` is lscpu's description of the CPU.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "paper-rsf-vesselseg_amd")
for _p in (ROOT, PKG):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "Gvoxel-iterations/s + % HBM roofline; batched vs plain per-block, same GPU"
# MI355X FP32 vector peak with one FMA per lane per cycle (packed FP32 is not
# generated): 256 CUs x 64 lanes x 2 FLOP x 2.4 GHz.
VALU_PEAK_TFLOPS = 78.6


def valu_roofline(ms_sweep1, voxels, literal):
    """Sweep 1's minimum arithmetic -- the separable 21-tap blur of each channel
    (2 channels, 4 in literal mode) along x, y and z, 63 FMA = 126 FLOP per
    channel-voxel -- over its mean launch time."""
    flop = (4 if literal else 2) * 3 * 21 * 2
    achieved = flop * voxels / (ms_sweep1 * 1e-3) / 1e12
    return {"kernel": "sweep1", "flop_per_voxel": flop, "achieved": round(achieved, 2),
            "peak": VALU_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": round(achieved / VALU_PEAK_TFLOPS, 4)}
HBM_PEAK_GBS = 8000.0
SWEEP_BYTES = {"sweep1": 9, "sweep2": 12}   # algorithmic bytes per voxel per launch


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def params(rsf_amd, iterations, mode):
    return rsf_amd.make_params(sigma=3.0, lambda1=1.0, nu=65.025, mu=1.0, dt=0.1, epsilon=1.0,
                               iterations=iterations, mode=mode)


def phantom(rsf_amd, n):
    img, _ = rsf_amd.synth_tubes((n, n, n), tubes=300, rmin=2.0, rmax=6.0, bias=20.0, noise=15.0, seed=2)
    phi0 = rsf_amd.initial_phi(img, 1.0, 110.0)
    log(f"generated {n}^3 phantom, {float((phi0 < 0).mean()):.3f} inside")
    return img, phi0


def plain_run(rsf_amd, torch, ctx, img, phi0, steps, warmup, mode, tune):
    n = img.shape[0]
    dev = torch.device("cuda:0")
    d_img = torch.from_numpy(img).to(dev)
    d_phi0 = torch.from_numpy(phi0).to(dev)
    d_phi = d_phi0.clone()
    blocks = [(0, n, n, n)]
    choice = None
    if tune:
        choice = ctx.autotune(blocks, rsf_amd.DTYPE_U8, params(rsf_amd, 1, mode))
        log(f"autotune: {choice}")
    if warmup > 0:
        ctx.run(d_img, d_phi, params(rsf_amd, warmup, mode))
    pk = params(rsf_amd, steps, mode)
    d_phi.copy_(d_phi0)
    ctx.run(d_img, d_phi, pk)          # captures and caches the graph of this loop
    d_phi.copy_(d_phi0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    loop_ms = ctx.run(d_img, d_phi, pk)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    vox = float(n) ** 3
    inside = float((d_phi < 0).float().mean().item())
    log(f"plain {n}^3 x {steps} ({'literal' if mode else 'fast'}): {wall * 1e3:.1f} ms wall, {loop_ms:.1f} ms loop")
    # per-kernel timing, eager launches with events around each sweep
    d_phi.copy_(d_phi0)
    ms1, ms2 = ctx.profile_sweeps(blocks, d_img, d_phi, params(rsf_amd, 20, mode))
    log(f"sweep1 {ms1:.3f} ms, sweep2 {ms2:.3f} ms per launch")
    return dict(wall=wall, loop_ms=loop_ms, vox=vox, ms1=ms1, ms2=ms2, choice=choice, img=img, phi0=phi0,
                inside=inside)


def batched_blocks(rsf_amd, nblocks=64):
    """The batched workload: nblocks uneven blocks, each side drawn from
    {96, 160, 192, 256} (numpy default_rng(3)), block b a tube phantom of seed
    1000+b with tubes in proportion to its volume (300 per 512^3, at least 4),
    phi0 the signed distance of the thresholded sigma-1 blur.  Returns (shapes
    (l, h, w), packed blocks, image, phi0) as flat numpy arrays."""
    import numpy as np

    rng = np.random.default_rng(3)
    sides = [96, 160, 192, 256]
    shapes = [tuple(int(s) for s in rng.choice(sides, 3)) for _ in range(nblocks)]
    blocks, total = rsf_amd.pack_blocks(shapes)
    img = np.empty(total, np.uint8)
    phi = np.empty(total, np.float32)
    for b, ((off, w, h, l), shp) in enumerate(zip(blocks, shapes)):
        nv = w * h * l
        tubes = max(4, int(round(300 * nv / 512 ** 3)))
        bi, _ = rsf_amd.synth_tubes(shp, tubes=tubes, rmin=2.0, rmax=6.0, bias=20.0, noise=15.0, seed=1000 + b)
        img[off:off + nv] = bi.ravel()
        phi[off:off + nv] = rsf_amd.initial_phi(bi, 1.0, 110.0).ravel()
    return shapes, blocks, img, phi


def batched_run(rsf_amd, torch, ctx, steps, mode, nblocks=64):
    shapes, blocks, img, phi = batched_blocks(rsf_amd, nblocks)
    total = img.size
    log(f"generated {nblocks} blocks, {total / 1e6:.1f} Mvoxel")
    dev = torch.device("cuda:0")
    d_img = torch.from_numpy(img).to(dev)
    d_phi0 = torch.from_numpy(phi).to(dev)
    p = params(rsf_amd, steps, mode)
    # the batched launch runs with the tile shapes the autotuner picks for this block set
    # (untimed); the per-block schedule takes each volume's defaults
    choice = ctx.autotune(blocks, rsf_amd.DTYPE_U8, params(rsf_amd, 1, mode))
    log(f"batched autotune: {choice}")
    out = {}
    for name, fn in (("batched", ctx.run_batched), ("per_block", ctx.run_sequential)):
        d_phi = d_phi0.clone()
        fn(blocks, d_img, d_phi, p)          # capture graphs
        d_phi.copy_(d_phi0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fn(blocks, d_img, d_phi, p)
        torch.cuda.synchronize()
        out[name] = time.perf_counter() - t0
        out[name + "_phi"] = d_phi
        log(f"{name}: {out[name] * 1e3:.1f} ms")
    same = bool(torch.equal(out["batched_phi"], out["per_block_phi"]))
    gv = lambda t: total * steps / t / 1e9  # noqa: E731
    return {
        "blocks": nblocks,
        "voxels": int(total),
        "iterations": steps,
        "value": round(gv(out["batched"]), 3),
        "plain_per_block": round(gv(out["per_block"]), 3),
        "speedup": round(out["per_block"] / out["batched"], 3),
        "kernels": choice,
        "results_identical": same,
    }


def large_run(rsf_amd, torch, ctx, steps, mode):
    """1024x1024x768 uint8 volume (2000 tubes, radius 1.5-8, noise sigma 20, seed 5), the
    default kernels and the autotuned kernels side by side."""
    l, h, w = 768, 1024, 1024
    t0 = time.perf_counter()
    img, _ = rsf_amd.synth_tubes((l, h, w), tubes=2000, rmin=1.5, rmax=8.0, bias=20.0, noise=20.0, seed=5)
    phi0 = rsf_amd.initial_phi(img, 1.0, 110.0)
    log(f"generated {w}x{h}x{l} phantom in {time.perf_counter() - t0:.1f} s")
    dev = torch.device("cuda:0")
    d_img = torch.from_numpy(img).to(dev)
    d_phi0 = torch.from_numpy(phi0).to(dev)
    del img, phi0
    d_phi = d_phi0.clone()
    blocks = [(0, w, h, l)]
    vox = float(w) * h * l
    out = {"volume": [w, h, l], "iterations": steps}
    for label in ("plain", "autotuned"):
        choice = None
        if label == "autotuned":
            choice = ctx.autotune(blocks, rsf_amd.DTYPE_U8, params(rsf_amd, 1, mode))
            log(f"large autotune: {choice}")
        d_phi.copy_(d_phi0)
        ctx.run(d_img, d_phi, params(rsf_amd, 2, mode))   # workspace, tile plans, image blur
        d_phi.copy_(d_phi0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        loop_ms = ctx.run(d_img, d_phi, params(rsf_amd, steps, mode))
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        out[label] = {"value": round(vox * steps / wall / 1e9, 3), "ms_per_iteration": round(wall * 1e3 / steps, 4),
                      "loop_ms_events": round(loop_ms, 1), "kernels": choice or "defaults"}
        log(f"large {label}: {wall:.2f} s, {out[label]['value']} Gvox-it/s")
    out["inside_fraction"] = round(float((d_phi < 0).float().mean().item()), 5)
    return out


def io_run(rsf_amd, torch, ctx):
    """Streaming .npy I/O at the large-volume size (1024x1024x768): a uint8 image and an
    fp32 phi file written to a temporary folder, then read into device memory by
    rsf_npy_load_device and by numpy.load + a torch copy, and phi written back by
    rsf_npy_save_device and by .cpu() + numpy.save.  The files sit in the page cache,
    so this measures the host-side path, not the disk."""
    import shutil
    import tempfile

    import numpy as np

    shape = (768, 1024, 1024)
    n = int(np.prod(shape))
    d = tempfile.mkdtemp()
    out = {"volume": [shape[2], shape[1], shape[0]]}
    try:
        rng = np.random.default_rng(7)
        np.save(os.path.join(d, "img.npy"), rng.integers(0, 256, shape, dtype=np.uint8))
        np.save(os.path.join(d, "phi.npy"), rng.normal(size=shape).astype(np.float32))
        for name, tdt, esz in (("img", torch.uint8, 1), ("phi", torch.float32, 4)):
            path = os.path.join(d, name + ".npy")
            dst = torch.empty(shape, dtype=tdt, device="cuda")
            ctx.npy_load(path, dst)   # warm the page cache and the staging allocations
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            ctx.npy_load(path, dst)
            torch.cuda.synchronize()
            ts = time.perf_counter() - t0
            t0 = time.perf_counter()
            ref = torch.from_numpy(np.load(path)).cuda()
            torch.cuda.synchronize()
            tn = time.perf_counter() - t0
            out["load_" + name] = {"gib": round(n * esz / 2**30, 3), "streamed_gbs": round(n * esz / ts / 1e9, 2),
                                   "numpy_gbs": round(n * esz / tn / 1e9, 2), "equal": bool(torch.equal(dst, ref))}
            del dst, ref
        src = torch.randn(shape, device="cuda")
        t0 = time.perf_counter()
        ctx.npy_save(os.path.join(d, "out.npy"), src)
        ts = time.perf_counter() - t0
        t0 = time.perf_counter()
        np.save(os.path.join(d, "out2.npy"), src.cpu().numpy())
        tn = time.perf_counter() - t0
        out["save_phi"] = {"streamed_gbs": round(4 * n / ts / 1e9, 2), "numpy_gbs": round(4 * n / tn / 1e9, 2)}
        del src
        # the whole rsf run file to file (rsf_run_npy, what the rsf executable calls): the
        # 512^3 phantom and its 0/255 seed streamed in, phi0 = sdf(seed - 13) on the device,
        # 100 literal iterations, phi streamed out
        shp = (512, 512, 512)
        img, mask = rsf_amd.synth_tubes(shp, tubes=300, seed=2)
        np.save(os.path.join(d, "e2e_img.npy"), img)
        np.save(os.path.join(d, "e2e_seed.npy"), (mask * 255).astype(np.uint8))
        del img, mask
        T = 100
        p = params(rsf_amd, T, rsf_amd.MODE_LITERAL)
        paths = [os.path.join(d, f) for f in ("e2e_img.npy", "e2e_seed.npy", "e2e_phi.npy")]
        ctx.run_npy(*paths, p)   # warm: graphs, workspace
        t0 = time.perf_counter()
        ms = ctx.run_npy(*paths, p, per_iteration=True)
        wall = time.perf_counter() - t0
        nv = 512 ** 3
        out["file_to_file"] = {"volume": [512, 512, 512], "iterations": T, "wall_s": round(wall, 3),
                               "loop_s": round(float(ms.sum()) / 1e3, 3),
                               "value": round(nv * T / wall / 1e9, 3), "unit": "Gvoxel-iterations/s",
                               "note": "includes .npy reads, PCIe, phi0 on the device and the .npy write"}
    finally:
        shutil.rmtree(d, ignore_errors=True)
    log(f"io: {out}")
    return out


def unet_run(rsf_amd, torch, ctx):
    """U-Net seed producer (rsf_amd.unet, random weights: the reference ships none) on a
    128x512x512 tube phantom: 16x64x64 tiles at stride 32, fp32, 64 tiles per forward,
    then the float64 vote; the (device) normalization is timed apart."""
    from rsf_amd import unet

    shape = (128, 512, 512)
    img, _ = rsf_amd.synth_tubes(shape, tubes=80, seed=9)
    torch.manual_seed(0)
    model = unet.UNet3D().cuda().eval()
    t0 = time.perf_counter()
    x = unet.normalize_device(img)
    torch.cuda.synchronize()
    t_norm = time.perf_counter() - t0
    sl = unet.tile_slices(shape)
    unet.predict_tiles(model, x, sl[:64], 64)   # warm up the convolution kernels
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    pred = unet.predict_tiles(model, x, sl, 64)
    torch.cuda.synchronize()
    t_pred = time.perf_counter() - t0
    t0 = time.perf_counter()
    prob, mask = unet.vote(pred, shape, sl)
    torch.cuda.synchronize()
    t_vote = time.perf_counter() - t0
    vox = float(shape[0] * shape[1] * shape[2])
    out = {"volume": [shape[2], shape[1], shape[0]], "tiles": len(sl), "normalize_ms": round(t_norm * 1e3, 1),
           "predict_ms": round(t_pred * 1e3, 1), "vote_ms": round(t_vote * 1e3, 1),
           "mvox_per_s": round(vox / (t_pred + t_vote) / 1e6, 1),
           "tiles_per_s": round(len(sl) / t_pred, 1), "inside_fraction": round(float(mask.double().mean()), 4)}
    log(f"unet: {out}")
    return out


def adjoint_run(rsf_amd, torch, ctx, K=32, C=8, nvol=8, n=128):
    """Differentiated batch: nvol volumes of n^3 (seeds 4000+b), K unrolled steps with
    checkpoints every C, BCE of sigmoid(-phi/eps) against the generator's masks; forward
    and backward timed separately."""
    import numpy as np

    from rsf_amd.autograd import bce_inside_loss, rsf_unrolled

    shapes = [(n, n, n)] * nvol
    blocks, total = rsf_amd.pack_blocks(shapes)
    img = np.empty(total, np.uint8)
    phi = np.empty(total, np.float32)
    mask = np.empty(total, np.float32)
    for b, (off, w, h, l) in enumerate(blocks):
        bi, bm = rsf_amd.synth_tubes((n, n, n), tubes=40, rmin=2.0, rmax=6.0, bias=20.0, noise=15.0, seed=4000 + b)
        img[off:off + n ** 3] = bi.ravel()
        mask[off:off + n ** 3] = bm.ravel()
        phi[off:off + n ** 3] = rsf_amd.initial_phi(bi, 1.0, 110.0).ravel()
    dev = torch.device("cuda:0")
    d_img = torch.from_numpy(img).to(dev)
    d_mask = torch.from_numpy(mask).to(dev)
    theta = torch.tensor([1.0, 1.0, 65.025, 1.0], device=dev, requires_grad=True)
    d_phi = torch.from_numpy(phi).to(dev).requires_grad_(True)
    times = {}
    for rep in range(2):   # the first repetition builds plans and workspaces
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = rsf_unrolled(d_phi, theta, d_img, blocks, K, C, sigma=3.0, epsilon=1.0, context=ctx)
        loss = bce_inside_loss(out, d_mask, 1.0)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        loss.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        times = {"forward_s": t1 - t0, "backward_s": t2 - t1}
        d_phi.grad = None
        theta.grad = None
    vs = total * K
    log(f"adjoint: forward {times['forward_s'] * 1e3:.1f} ms, backward {times['backward_s'] * 1e3:.1f} ms")
    return {
        "volumes": nvol, "shape": [n, n, n], "unrolled_steps": K, "checkpoint_every": C,
        "forward_ms": round(times["forward_s"] * 1e3, 2), "backward_ms": round(times["backward_s"] * 1e3, 2),
        "forward_Gvox_steps_per_s": round(vs / times["forward_s"] / 1e9, 3),
        "backward_Gvox_steps_per_s": round(vs / times["backward_s"] / 1e9, 3),
        "loss": round(float(loss.detach()), 6),
    }


def host_cpu():
    """The host's CPU as lscpu reports it (model name, sockets, cores), read before
    anything touches the GPU; /proc/cpuinfo fields as a fallback."""
    import subprocess

    info = {"nproc": os.cpu_count()}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for ln in out.splitlines():
            k, _, v = ln.partition(":")
            if k.strip() in ("Model name", "Vendor ID", "CPU family", "Model", "Stepping", "Socket(s)",
                             "Core(s) per socket", "Thread(s) per core", "CPU max MHz", "L3 cache"):
                info[k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    if "Model name" not in info:
        try:
            with open("/proc/cpuinfo") as f:
                info["Model name"] = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")),
                                          "")
        except OSError:
            pass
    info["model_name"] = info.get("Model name", "")
    return info


def config1_cpu(rsf_amd, threads):
    """The CPU configuration run in full: 128^3 uint8 (40 tubes, seed 1), sigma 3, lambda 1,
    nu 65.025, mu 1, dt 0.1, eps 1, 100 iterations of the oracle on `threads` host threads."""
    import numpy as np

    import oracle

    img, _ = rsf_amd.synth_tubes((128, 128, 128), tubes=40, rmin=2.0, rmax=6.0, bias=20.0, noise=15.0, seed=1)
    phi0 = rsf_amd.initial_phi(img, 1.0, 110.0)
    p = oracle.make_params(sigma=3.0, lambda1=1.0, nu=65.025, mu=1.0, dt=0.1, epsilon=1.0)
    t0 = time.perf_counter()
    phi = oracle.run(phi0, img.astype(np.float32), p, 100, threads)
    t = time.perf_counter() - t0
    log(f"cpu oracle, config 1: 100 iterations of 128^3 in {t:.1f} s on {threads} threads")
    return {"workload": "128^3 uint8, 40 tubes (seed 1), 100 iterations, run in full", "seconds": round(t, 2),
            "value": round(128 ** 3 * 100 / t / 1e9, 5), "unit": "Gvoxel-iterations/s", "cores": threads,
            "inside_fraction": round(float((phi < 0).mean()), 5)}


def cpu_baseline(img, phi0, budget_s, rsf_amd=None, host=None):
    import numpy as np

    import oracle

    threads = min(16, os.cpu_count() or 1)
    p = oracle.make_params(sigma=3.0, lambda1=1.0, nu=65.025, mu=1.0, dt=0.1, epsilon=1.0)
    I = img.astype(np.float32)
    t0 = time.perf_counter()
    phi = oracle.run(phi0, I, p, 1, threads)
    t1 = time.perf_counter() - t0
    iters = 1
    extra = int(max(0, min(8, budget_s / max(t1, 1e-3) - 1)))
    if extra:
        t0 = time.perf_counter()
        oracle.run(phi, I, p, extra, threads)
        t1 += time.perf_counter() - t0
        iters += extra
    n = phi0.size
    log(f"cpu oracle: {iters} iterations of {n / 1e6:.0f} Mvoxel in {t1:.1f} s on {threads} threads")
    # one thread, on a bounded crop of the same volume (one iteration of at most 256^3)
    c = tuple(min(256, d) for d in phi0.shape)
    cphi = np.ascontiguousarray(phi0[:c[0], :c[1], :c[2]])
    cI = np.ascontiguousarray(I[:c[0], :c[1], :c[2]])
    t0 = time.perf_counter()
    oracle.run(cphi, cI, p, 1, 1)
    t_one = time.perf_counter() - t0
    log(f"cpu oracle, 1 thread: 1 iteration of {c[2]}x{c[1]}x{c[0]} in {t_one:.1f} s")
    host = host or host_cpu()
    out = {
        "value": round(n * iters / t1 / 1e9, 5),
        "unit": "Gvoxel-iterations/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{iters} iteration(s) of the same {round(n ** (1 / 3))}^3 workload (oracle: unfused FP32 "
                  f"restatement of CUDA_RSF/main.cpp:234-300, OpenMP)",
        "one_thread": {"value": round(cphi.size / t_one / 1e9, 5), "unit": "Gvoxel-iterations/s",
                       "sample": f"1 iteration of a {c[2]}x{c[1]}x{c[0]} crop of the same workload"},
        "host": host,
    }
    if rsf_amd is not None:
        out["config1_full_run"] = config1_cpu(rsf_amd, threads)
    return out


def traffic_from_profiles(kernel, mode):
    """HBM bytes per launch of a sweep measured by rocprofv3 (profiles/traffic.json)."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            return json.load(f).get(mode, {}).get(kernel, {}).get("bytes_per_launch")
    except (OSError, ValueError):
        return None


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--size", type=int, default=512, help="cube side of the plain workload")
    ap.add_argument("--mode", choices=["fast", "literal"], default="literal")
    ap.add_argument("--no-fast", action="store_true", help="skip the fast-mode line of the literal run")
    ap.add_argument("--tune", action="store_true",
                    help="autotune the plain runs too (the plain drop-in runs the library defaults)")
    ap.add_argument("--no-tune", action="store_true", help="the default for the plain runs (kept for old commands)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--batched-steps", type=int, default=300)
    ap.add_argument("--cpu-budget", type=float, default=20.0, help="seconds of CPU baseline work")
    ap.add_argument("--no-large", action="store_true",
                    help="skip the 1024x1024x768 volume (2000 tubes), default and autotuned kernels")
    ap.add_argument("--large", action="store_true", help="(the default; kept for old commands)")
    ap.add_argument("--large-steps", type=int, default=None, help="iterations of the large volume (default --steps)")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip the differentiated batch (8 x 128^3, 32 steps, checkpoints every 8)")
    ap.add_argument("--adjoint", action="store_true", help="(the default; kept for old commands)")
    ap.add_argument("--unet", action="store_true",
                    help="also time the U-Net seed producer's tiled inference on 128x512x512")
    ap.add_argument("--io", action="store_true",
                    help="also time streaming .npy load/save of a 1024x1024x768 volume against numpy")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run 5 device-to-device copies of MIB MiB (scale for rocprofv3 byte counters)")
    args = ap.parse_args()
    host = host_cpu() if not args.no_cpu else None   # lscpu runs before the GPU is initialised
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: the RSF engine runs each volume or batch of blocks on a single "
              "device; sharding over several GPUs is not implemented", file=sys.stderr)
        sys.exit(2)

    import torch

    import rsf_amd

    if not torch.cuda.is_available():
        print("bench.py needs a HIP device", file=sys.stderr)
        sys.exit(1)
    if args.copy_calibration > 0:
        n = args.copy_calibration << 18
        a = torch.arange(n, dtype=torch.float32, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": 4 * n, "copies": 5, "ok": bool(torch.equal(a, b))}))
        return
    rsf_amd.lib()
    mode = rsf_amd.MODE_FAST if args.mode == "fast" else rsf_amd.MODE_LITERAL
    ctx = rsf_amd.Context(0)
    img, phi0 = phantom(rsf_amd, args.size)
    r = plain_run(rsf_amd, torch, ctx, img, phi0, args.steps, args.warmup, mode, args.tune and not args.no_tune)
    N, K = r["vox"], args.steps
    value = N * K / r["wall"] / 1e9
    dom = "sweep1" if r["ms1"] >= r["ms2"] else "sweep2"
    ms_dom = r["ms1"] if dom == "sweep1" else r["ms2"]
    achieved = SWEEP_BYTES[dom] * N / (ms_dom * 1e-3) / 1e9
    traffic = traffic_from_profiles(dom, args.mode)
    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Gvoxel-iterations/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": round(r["wall"] * 1e3 / K, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"plain RSF drop-in: {args.size}^3 uint8 tube phantom (300 tubes, seed 2), fp32 phi, "
                        f"sigma 3 (21 taps), lambda 1, nu 65.025, mu 1, dt 0.1, eps 1, {K} iterations, "
                        f"{args.mode} mode",
            "volume": [args.size] * 3,
            "iterations": K,
            "mode": args.mode,
            "kernels": r["choice"],
            "loop_ms_events": round(r["loop_ms"], 3),
        },
        "roofline": {
            "kernel": dom,
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "algorithmic_bytes_per_voxel": SWEEP_BYTES[dom],
            "ms_per_launch": {"sweep1": round(r["ms1"], 4), "sweep2": round(r["ms2"], 4)},
            # sweep 1 is bounded by FP32 FMA issue, not by HBM: its minimum arithmetic
            # against the VALU peak without packed FP32
            "valu": valu_roofline(r["ms1"], N, mode == rsf_amd.MODE_LITERAL),
        },
        "loop_roofline": {
            "bytes_per_voxel_iteration": 21,
            "achieved": round(21 * N * K / r["wall"] / 1e9, 1),
            "frac": round(21 * N * K / r["wall"] / 1e9 / HBM_PEAK_GBS, 4),
        },
    }
    if not args.no_batched:
        line["batched"] = batched_run(rsf_amd, torch, ctx, args.batched_steps, mode)
    if mode == rsf_amd.MODE_LITERAL and not args.no_fast:
        rf = plain_run(rsf_amd, torch, ctx, img, phi0, args.steps, args.warmup, rsf_amd.MODE_FAST,
                       args.tune and not args.no_tune)
        line["fast_mode"] = {
            "value": round(N * K / rf["wall"] / 1e9, 3),
            "ms_per_step": round(rf["wall"] * 1e3 / K, 4),
            "ms_per_launch": {"sweep1": round(rf["ms1"], 4), "sweep2": round(rf["ms2"], 4)},
            "valu": valu_roofline(rf["ms1"], N, False),
            "kernels": rf["choice"],
            "parity": "one-step gate met (max|dphi| <= 2e-5 max|phi|); full-run sign masks within FP32 "
                      "noise, not bit-identical",
        }
    if not args.no_adjoint:
        line["adjoint"] = adjoint_run(rsf_amd, torch, ctx)
    if not args.no_large:
        line["large"] = large_run(rsf_amd, torch, ctx, args.large_steps or args.steps, mode)
    if args.io:
        line["io"] = io_run(rsf_amd, torch, ctx)
    if args.unet:
        line["unet"] = unet_run(rsf_amd, torch, ctx)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(r["img"], r["phi0"], args.cpu_budget, rsf_amd, host)
    ctx.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
