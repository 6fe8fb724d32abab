"""Benchmark: forward+backward rasterisation throughput on one MI355X.

Workload (default): 3,000,000 SH-degree-3 Gaussians in a 10x10x10 box (20% in
50 blobs), one 1920x1080 camera, L1 loss against a seeded target.  One step =
fused preprocess + binning + sort + forward compositing + loss gradient +
reverse-order backward compositing + projection adjoint, i.e. everything
between raw Gaussian parameters and their gradients.  Inputs are resident in
HBM before timing starts.

Prints one JSON line (metric/value/unit per the workload definition) with a
`roofline` object for the dominant kernel (algorithmic bytes of the
(Gaussian, tile) pairs the launch processes / HIP-event time of that kernel
on the stream it runs on, against 8 TB/s; beside it the DRAM bytes the
hardware counters measure and the VALU issue rate) and a `cpu_baseline`
object (the CPU oracle on all host cores, on a stratified 5% tile sample of
the same workload, scaled, median of 3).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np
import torch

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # B/s, MI355X HBM3E
WORKLOADS = ("dense_3m", "fwd_bwd_500k", "plain_forward_10k", "multiview_6m", "serve", "train_step_3m")


def make_scene(name):
    from gaussian_splatting_3d_amd import scenes

    return {"dense_3m": scenes.dense_3m, "fwd_bwd_500k": scenes.fwd_bwd_500k,
            "plain_forward_10k": scenes.plain_forward_10k, "multiview_6m": scenes.multiview_6m}[name]()


def alg_bytes(P, n_pix, C, V, N):
    """Algorithmic HBM bytes per launch, per stage (FP32); n_pix = pixels of
    all views."""
    K = 3 * C * C
    return {
        "preprocess": 44 * N + 36 * V,
        "composite_fwd": P * (28 + 4 * K) + 20 * n_pix,
        "composite_bwd": P * (28 + 4 * K) + P * 8 * (6 + K) + 32 * n_pix,
        "project_bwd": V * 24 + 44 * V + 44 * N,
    }


class Timed:
    """HIP events around a region on the current stream; averaged."""

    def __init__(self):
        self.pairs = []

    def __enter__(self):
        self.s = torch.cuda.Event(enable_timing=True)
        self.e = torch.cuda.Event(enable_timing=True)
        self.s.record()
        return self

    def __exit__(self, *a):
        self.e.record()
        self.pairs.append((self.s, self.e))

    def mean_ms(self):
        return float(np.mean([s.elapsed_time(e) for s, e in self.pairs])) if self.pairs else float("nan")


def examined_pairs(n_done, H, W, tile=16):
    """(Gaussian, tile) pairs a compositing launch processes: per tile, the
    list entries up to where its last pixel saturated (the largest n_done of
    its pixels), summed over the tiles."""
    nth, ntw = -(-H // tile), -(-W // tile)
    nd = torch.zeros(nth * tile, ntw * tile, dtype=torch.int64, device=n_done.device)
    nd[:H, :W] = n_done.view(H, W).to(torch.int64)
    return int(nd.view(nth, tile, ntw, tile).amax(dim=(1, 3)).sum())


def gpu_bench(scene, steps, warmup, exact=False, split=False, cmf=False):
    from gaussian_splatting_3d_amd import _lib, rasterize, scenes, training
    from gaussian_splatting_3d_amd.backend import _backend

    dev = torch.device("cuda:0")
    c2w, cam = scene.cameras[0]
    cfg = scenes.RasterConfig()
    params = [t.to(dev).contiguous() for t in (scene.mean, scene.qvec, scene.svec_before_activation,
                                               scene.alpha_before_activation, scene.sh_coeffs)]
    target = scenes.target_image(cam.h, cam.w, 3).to(dev).reshape(-1)
    t_fwd, t_bwd = Timed(), Timed()
    orig_fwd, orig_bwd, orig_fwd2 = _backend.sh_forward, _backend.sh_backward, _backend.sh_forward_two_level
    state = {"on": False}

    def fwd_wrap(*a, **k):
        if not state["on"]:
            return orig_fwd(*a, **k)
        with t_fwd:
            return orig_fwd(*a, **k)

    def fwd2_wrap(*a, **k):
        if not state["on"]:
            return orig_fwd2(*a, **k)
        with t_fwd:
            return orig_fwd2(*a, **k)

    def bwd_wrap(*a, **k):
        if not state["on"]:
            return orig_bwd(*a, **k)
        with t_bwd:
            return orig_bwd(*a, **k)

    _backend.sh_forward, _backend.sh_backward, _backend.sh_forward_two_level = fwd_wrap, bwd_wrap, fwd2_wrap

    def step():
        out, st = rasterize.forward_view(*params, c2w, cam, scene.C, cfg, exact=exact, prepare_backward=True)
        # mean |out - target| and its gradient sign(out - target) / numel, one fused pass
        _, grad_out = training.photometric_loss(out, target.view_as(out), "l1")
        grads = rasterize.backward_view(st, *params, grad_out, split_bf16=split, colour_mfma=cmf)
        return st, grads

    try:
        for _ in range(warmup):
            step()
        torch.cuda.synchronize()
        state["on"] = True
        t0 = time.perf_counter()
        for _ in range(steps):
            st, grads = step()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        state["on"] = False
        # forward-only throughput (image only, no backward state kept)
        torch.cuda.synchronize()
        f0 = time.perf_counter()
        for _ in range(steps):
            rasterize.forward_view(*params, c2w, cam, scene.C, cfg, exact=exact)
        torch.cuda.synchronize()
        f1 = time.perf_counter()
    finally:
        _backend.sh_forward, _backend.sh_backward = orig_fwd, orig_bwd
        _backend.sh_forward_two_level = orig_fwd2
    b = st.b
    for g in grads[:5]:
        assert torch.isfinite(g).all()
    # (Gaussian, 16-px tile) pairs, the unit of the compositing bytes; the
    # two-level forward bins per 64-px block, so count them once, untimed
    P = rasterize.preprocess_and_bin(*params[:4], c2w, cam, D=cfg.tile_culling_radius).P if b.shift else b.P
    return dict(ms_step=(t1 - t0) * 1e3 / steps, ms_fwd_only=(f1 - f0) * 1e3 / steps,
                ms_composite_fwd=t_fwd.mean_ms(), ms_composite_bwd=t_bwd.mean_ms(), N=b.N, V=b.V, P=P,
                P_exam=examined_pairs(st.n_done, cam.h, cam.w), block_pairs=b.P if b.shift else None, H=cam.h,
                W=cam.w, C=scene.C)


def gpu_bench_views(scene, steps, warmup):
    """One training step over all cameras of the scene in one batch: batched
    forward, summed L1 against seeded targets, batched adjoint.  HIP events
    bracket the two compositing launches (on the stream they run on)."""
    from gaussian_splatting_3d_amd import _lib, multiview, scenes, training

    dev = torch.device("cuda:0")
    cfg = scenes.RasterConfig()
    params = [t.to(dev).contiguous() for t in (scene.mean, scene.qvec, scene.svec_before_activation,
                                               scene.alpha_before_activation, scene.sh_coeffs)]
    g = torch.Generator(device=dev).manual_seed(105)
    n_pix = sum(cam.h * cam.w for _, cam in scene.cameras)
    target = torch.rand(n_pix * 3, generator=g, device=dev)
    L = _lib.load()
    t_fwd, t_bwd = Timed(), Timed()
    state = {"on": False}
    names = ("gsmi_render_views_sh", "gsmi_render_views_sh_two_level_record", "gsmi_render_views_backward_sh",
             "gsmi_render_views_backward_sh_lists")
    orig = {k: getattr(L, k) for k in names}

    def wrap(name, timer):
        def f(*a):
            if not state["on"]:
                return orig[name](*a)
            with timer:
                return orig[name](*a)
        return f

    for k in names:  # the per-tile or the two-level entries, whichever the step calls
        setattr(L, k, wrap(k, t_fwd if "backward" not in k else t_bwd))

    def step():
        out, st = multiview.forward_views(*params, scene.cameras, scene.C, cfg)
        _, grad_out = training.photometric_loss(out, target.view_as(out), "l1")
        grads = multiview.backward_views(st, *params, grad_out)
        return st, grads

    try:
        for _ in range(warmup):
            step()
        torch.cuda.synchronize()
        state["on"] = True
        t0 = time.perf_counter()
        for _ in range(steps):
            st, grads = step()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        state["on"] = False
    finally:
        for k, v in orig.items():
            setattr(L, k, v)
    for gr in grads:
        assert torch.isfinite(gr).all()
    b = st.b
    P_exam = sum(examined_pairs(st.n_done[m["pix_base"]:m["pix_base"] + m["H"] * m["W"]], m["H"], m["W"])
                 for m in b.meta)
    return dict(ms_step=(t1 - t0) * 1e3 / steps, ms_composite_fwd=t_fwd.mean_ms(), ms_composite_bwd=t_bwd.mean_ms(),
                N=b.N, V=b.V, P=b.P, shift=b.shift, P_exam=P_exam, n_pix=n_pix, views=b.B, C=scene.C,
                H=scene.cameras[0][1].h, W=scene.cameras[0][1].w)


def host_threads():
    """CPU threads this process may use: OMP_NUM_THREADS when set (the share
    of the host a GPU job gets), else the CPUs it is allowed to run on."""
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    return len(os.sched_getaffinity(0))


def cpu_model():
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(scene, reps=3, tile_frac=0.05):
    """The CPU oracle on all host threads, on the same seeded workload:
    full preprocessing (activation, cull, projection, tile bounds) and
    binning (key emission + stable sort + ranges); compositing forward and
    backward on a stratified sample of tile_frac of the non-empty tiles
    (every k-th tile in order of list length, so short and long lists are
    both represented), scaled by non-empty / sampled tiles; projection
    adjoint on a sample of the visible Gaussians, scaled by V.  Median of
    `reps` repetitions of the extrapolated step time."""
    from oracle import oracle as O

    threads = host_threads()
    torch.set_num_threads(threads)
    O.set_threads(threads)
    c2w, cam = scene.cameras[0]
    ocam = O.Camera(cam.fx, cam.fy, cam.cx, cam.cy, cam.w, cam.h, cam.near_plane, cam.far_plane)
    C = scene.C
    mean = scene.mean.numpy()
    steps = []
    for rep in range(reps):
        t0 = time.perf_counter()
        s, a = O.activate(scene.svec_before_activation, scene.alpha_before_activation)
        normals, pts = O.frustum(c2w, ocam)
        mask = O.cull(scene.mean, s, normals, pts, 1.0)
        idx = np.nonzero(mask)[0]
        m2d, cov, depth = O.project(mean[idx], scene.qvec.numpy()[idx], s[idx], c2w)
        P, tl, br = O.aabb(m2d, cov, ocam, 6.0)
        ic = O.image_constants(ocam)
        bn = O.bin_gaussians(tl, br, depth, ic["n_tiles_h"], ic["n_tiles_w"])
        t_pre = time.perf_counter() - t0
        n_tiles = ic["n_tiles_h"] * ic["n_tiles_w"]
        counts = np.where(bn["start"] >= 0, bn["end"] - bn["start"], 0).astype(np.int64)
        nonempty = np.nonzero(counts)[0]
        by_len = nonempty[np.argsort(counts[nonempty], kind="stable")]
        k = max(1, int(round(1.0 / tile_frac)))
        chosen = by_len[k // 2::k]
        st = -np.ones(n_tiles, np.int32)
        en = -np.ones(n_tiles, np.int32)
        st[chosen] = bn["start"][chosen]
        en[chosen] = bn["end"][chosen]
        sampled_pairs = int(counts[chosen].sum())
        sh_v = np.ascontiguousarray(scene.sh_coeffs.numpy()[idx][..., : C * C])
        al = a[idx]
        c0 = time.perf_counter()
        out, T, nd, _ = O.render_sh_fwd(m2d, cov, sh_v, al, st, en, bn["ids"], ic["topleft"], c2w, 16,
                                        ic["n_tiles_h"], ic["n_tiles_w"], ic["psx"], ic["psy"], ic["H"], ic["W"], C,
                                        1e-4)
        gout = np.sign(out - 0.5).astype(np.float32) / out.size
        O.render_sh_bwd(m2d, cov, sh_v, al, st, en, bn["ids"], out, gout, ic["topleft"], c2w, 16, ic["n_tiles_h"],
                        ic["n_tiles_w"], ic["psx"], ic["psy"], ic["H"], ic["W"], C, 1e-4)
        t_comp = time.perf_counter() - c0
        scale = nonempty.size / max(chosen.size, 1)
        ns = min(len(idx), 100_000)
        sub = idx[:ns]
        a0 = time.perf_counter()
        O.projection_adjoint(mean[sub], scene.qvec.numpy()[sub], scene.svec_before_activation.numpy()[sub],
                             scene.alpha_before_activation.numpy()[sub], scene.sh_coeffs.numpy()[sub], c2w,
                             np.arange(ns), np.zeros((ns, 2), np.float32), np.zeros((ns, 4), np.float32),
                             np.zeros((ns, 3, C * C), np.float32), np.zeros(ns, np.float32), C)
        t_adj = (time.perf_counter() - a0) * len(idx) / max(ns, 1)
        steps.append((t_pre + t_comp * scale + t_adj, t_pre, t_comp, t_adj))
        print(f"[bench] cpu baseline rep {rep}: step {steps[-1][0]:.1f}s (preprocess+binning {t_pre:.1f}s, "
              f"compositing sample {t_comp:.2f}s x{scale:.1f}, adjoint {t_adj:.1f}s)", file=sys.stderr, flush=True)
    med = sorted(steps)[len(steps) // 2]
    t_step = med[0]
    return dict(value=ic["H"] * ic["W"] / t_step / 1e6, unit="Mpix/s", cores=threads, kind="port",
                cpu_model=cpu_model(), step_s=round(t_step, 2),
                sample=(f"CPU oracle on {threads} threads ({cpu_model()}), median of {reps}: full "
                        f"preprocess+binning of N={scene.N} ({med[1]:.1f}s); fwd+bwd compositing on {chosen.size} "
                        f"of {nonempty.size} non-empty tiles (every {k}th by list length, {sampled_pairs} of {P} "
                        f"pairs, {med[2]:.2f}s, scaled x{scale:.1f}); projection adjoint on {ns} of {len(idx)} "
                        f"visible Gaussians ({med[3]:.1f}s scaled); extrapolated step {t_step:.1f}s"))


KERNEL_PREFIX = {"composite_fwd": "gsmi::k_fwd_sh", "composite_bwd": "gsmi::k_bwd_sh"}


def _pmc_pass(counters, child_args, timeout_s=90):
    """One rocprofv3 counter pass (counters: a name or a list of names that
    fit one pass) over a child run of this script; returns {kernel name:
    {counter: mean value per dispatch}} or raises."""
    import shutil
    import signal
    import sqlite3
    import subprocess
    import tempfile

    counters = [counters] if isinstance(counters, str) else list(counters)
    exe = shutil.which("rocprofv3")
    if exe is None:
        raise RuntimeError("rocprofv3 not found")
    out = tempfile.mkdtemp(prefix="gsmi_pmc_", dir="/tmp")
    cmd = [exe, "--pmc", *counters, "-d", out, "-o", "run", "--", sys.executable, os.path.abspath(__file__)] + \
        child_args
    env = dict(os.environ, TMPDIR="/tmp")
    proc = subprocess.Popen(cmd, cwd="/tmp", env=env, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL,
                            start_new_session=True)
    try:
        rc = proc.wait(timeout=timeout_s)
    except subprocess.TimeoutExpired:
        os.killpg(proc.pid, signal.SIGKILL)
        proc.wait()
        raise RuntimeError(f"rocprofv3 --pmc {counters} timed out")
    db = os.path.join(out, "run_results.db")
    if not os.path.exists(db):
        raise RuntimeError(f"rocprofv3 --pmc {counters} exited {rc} without results")
    con = sqlite3.connect(db)
    rows = con.execute("select kernel_name, counter_name, avg(value) from counters_collection "
                       "group by kernel_name, counter_name").fetchall()
    con.close()
    shutil.rmtree(out, ignore_errors=True)
    res = {}
    for name, counter, v in rows:
        res.setdefault(name, {})[counter] = float(v)
    return res


def _kernel_values(run, kernel):
    """The counter values of the first kernel whose name starts with the
    dominant kernel's prefix."""
    for k, vals in run.items():
        if k.split("void ", 1)[-1].startswith(KERNEL_PREFIX[kernel]):
            return vals
    raise RuntimeError(f"{kernel} not seen in the counter run")


# Nominal non-packed FP32 VALU issue rate of MI355X: 256 CUs x 4 SIMDs x one
# wave64 instruction per 4 cycles at the 2.4 GHz peak engine clock.  Some
# VALU instructions (moves, lane reads, compares) take fewer cycles, so a
# kernel rich in them can measure above it.
VALU_PEAK_INSTR_S = 256 * 4 * 2.4e9 / 4


def _child_args(workload, exact, bwd_opts=()):
    return ["--pmc-child", "--workload", workload] + (["--exact"] if exact else []) + list(bwd_opts)


def measure_valu(workload, exact, kernel, kernel_ms, bwd_opts=()):
    """Wave-level VALU instructions per launch of `kernel` (rocprofv3
    SQ_INSTS_VALU, summed over the chip) and their issue rate over the
    kernel's HIP-event time, against the non-packed issue peak."""
    vals = _kernel_values(_pmc_pass("SQ_INSTS_VALU", _child_args(workload, exact, bwd_opts)), kernel)
    n = vals["SQ_INSTS_VALU"]
    rate = n / (kernel_ms / 1e3)
    return dict(instructions_per_launch=int(n), achieved_g_instr_s=round(rate / 1e9, 2),
                peak_g_instr_s=round(VALU_PEAK_INSTR_S / 1e9, 1), frac=round(rate / VALU_PEAK_INSTR_S, 4))


FP32_PEAK_FLOP_S = 157.3e12  # MI355X FP32 vector peak (published)


def measure_flops(workload, exact, kernel, kernel_ms, bwd_opts=()):
    """FP32 floating-point operations per launch of `kernel` from one
    counter pass, as rocprofv3's TOTAL_32_OPS metric counts them without the
    integer term: 64 lanes x (2 SQ_INSTS_VALU_FMA_F32 + SQ_INSTS_VALU_ADD_F32 +
    SQ_INSTS_VALU_MUL_F32 + SQ_INSTS_VALU_TRANS_F32) + 512 x
    SQ_INSTS_VALU_MFMA_MOPS_F32 (every lane counted, active or not), over the
    kernel's HIP-event time, against the 157.3 TFLOP/s FP32 peak."""
    names = ["SQ_INSTS_VALU_FMA_F32", "SQ_INSTS_VALU_ADD_F32", "SQ_INSTS_VALU_MUL_F32", "SQ_INSTS_VALU_TRANS_F32",
             "SQ_INSTS_VALU_MFMA_MOPS_F32"]
    vals = _kernel_values(_pmc_pass(names, _child_args(workload, exact, bwd_opts)), kernel)
    fma, add, mul, trans, mops = (vals.get(k, 0.0) for k in names)
    vec = 64.0 * (2.0 * fma + add + mul + trans)
    mfma = 512.0 * mops
    rate = (vec + mfma) / (kernel_ms / 1e3)
    return dict(flops_per_launch=int(vec + mfma), vector_flops=int(vec), mfma_flops=int(mfma),
                wave_instr={"fma": int(fma), "add": int(add), "mul": int(mul), "trans": int(trans)},
                achieved_tflop_s=round(rate / 1e12, 3), peak_tflop_s=FP32_PEAK_FLOP_S / 1e12,
                frac=round(rate / FP32_PEAK_FLOP_S, 4))


def measure_traffic(workload, exact, kernel, bwd_opts=()):
    """HBM bytes per launch of `kernel` from rocprofv3 FETCH_SIZE and
    WRITE_SIZE (KiB per dispatch), each scaled by what the same counter reads
    for a 1 GiB device-to-device copy (bytes moved / counter KiB)."""
    copy_bytes = float(1 << 30)
    res = {}
    child = _child_args(workload, exact, bwd_opts)
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        cal = _pmc_pass(counter, ["--copy-calibration"])
        copy_kib = [v[counter] for k, v in cal.items() if "copyBuffer" in k and counter in v]
        if not copy_kib or copy_kib[0] <= 0:
            raise RuntimeError(f"no copy kernel seen in the {counter} calibration run")
        scale = copy_bytes / copy_kib[0]  # bytes per counter KiB
        res[counter] = (_kernel_values(_pmc_pass(counter, child), kernel)[counter], scale)
    fetch = res["FETCH_SIZE"][0] * res["FETCH_SIZE"][1]
    write = res["WRITE_SIZE"][0] * res["WRITE_SIZE"][1]
    return dict(bytes=fetch + write, fetch_bytes=fetch, write_bytes=write,
                fetch_scale_bytes_per_kib=res["FETCH_SIZE"][1], write_scale_bytes_per_kib=res["WRITE_SIZE"][1])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="dense_3m", choices=WORKLOADS)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--exact", action="store_true", help="exact colour arithmetic (bit-reproducible images)")
    ap.add_argument("--bwd-split", action="store_true",
                    help="opt-in: SH-gradient products on bf16 matrix cores as a three-term split (not fp32; the "
                         "line's dtype says so)")
    ap.add_argument("--bwd-colour-mfma", action="store_true",
                    help="opt-in: the backward's colours on fp32 matrix cores (five entries at a time)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device-to-device copy 3 times (scales rocprof byte counters)")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the two rocprofv3 counter passes that measure the dominant kernel's HBM bytes")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving measurement of the default run")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the 8-view 4K step and the training-iteration measurements of the default run")
    ap.add_argument("--serve-levels", default="0.7,0.85,1,1.15,1.3,1.45,1.6,1.8,2,2.3,2.6",
                    help="server arrival rates as multiples of the plain per-call capacity")
    ap.add_argument("--serve-duration", type=float, default=2.0, help="seconds of arrivals per serving level")
    ap.add_argument("--pmc-child", action="store_true", help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.copy_calibration:
        src = torch.empty(1 << 28, dtype=torch.float32, device="cuda:0").fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read_per_launch": src.numel() * 4, "copy_bytes_written_per_launch": dst.numel() * 4}))
        return
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)
    if int(os.environ.get("RANK", "0")) != 0:
        return
    if args.pmc_child:  # one step under a counter pass of measure_traffic
        sc = make_scene(args.workload)
        if len(sc.cameras) > 1:
            gpu_bench_views(sc, 1, 1)
        else:
            gpu_bench(sc, 1, 1, exact=args.exact, split=args.bwd_split, cmf=args.bwd_colour_mfma)
        return
    if args.workload == "serve":
        serve_main(args)
        return
    if args.workload == "train_step_3m":
        train_main(args)
        return
    scene = make_scene(args.workload)
    bwd_opts = (["--bwd-split"] if args.bwd_split else []) + (["--bwd-colour-mfma"] if args.bwd_colour_mfma else [])
    print(f"[bench] workload {scene.name}: N={scene.N}", file=sys.stderr, flush=True)
    views = len(scene.cameras)
    if views > 1:
        r = gpu_bench_views(scene, args.steps, args.warmup)
    else:
        r = gpu_bench(scene, args.steps, args.warmup, exact=args.exact, split=args.bwd_split, cmf=args.bwd_colour_mfma)
        r["n_pix"], r["views"] = r["H"] * r["W"], 1
    print(f"[bench] gpu done: {r}", file=sys.stderr, flush=True)
    # compositing bytes are counted over the pairs the launches process (the
    # list entries up to each tile's saturation), not over every listed pair
    ab = alg_bytes(r["P_exam"], r["n_pix"], r["C"], r["V"], r["N"])
    dom = "composite_bwd" if r["ms_composite_bwd"] >= r["ms_composite_fwd"] else "composite_fwd"
    t_dom = r["ms_" + dom] / 1e3
    achieved = ab[dom] / t_dom / 1e9
    mpix = r["n_pix"] / (r["ms_step"] / 1e3) / 1e6
    metric = {"dense_3m": "fwd+bwd Mpix/s @3M Gaussians 1080p",
              "multiview_6m": "fwd+bwd Mpix/s @6M Gaussians, 8 views 3840x2160"}.get(
        args.workload, f"fwd+bwd Mpix/s ({args.workload})")
    line = {
        "metric": metric,
        "value": round(mpix, 3),
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["ms_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+bf16x3-split" if args.bwd_split else "f32",
        "data": "synthetic (seeded scene and target, random-init Gaussians)",
        "config": {"workload": scene.name, "gaussians": r["N"], "visible": r["V"],
                   ("block_pairs" if r.get("shift") else "pairs"): r["P"],
                   "width": r["W"], "height": r["H"], "sh_degree": r["C"] - 1, "views": r["views"]},
        "gaussians_per_s_fwd_bwd": round(r["N"] / (r["ms_step"] / 1e3), 1),
        "kernels_ms": {"composite_fwd": round(r["ms_composite_fwd"], 4),
                       "composite_bwd": round(r["ms_composite_bwd"], 4)},
        "roofline": {"kernel": dom, "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": round(achieved * 1e9 / HBM_PEAK, 4), "traffic": None,
                     "alg_bytes_per_launch": int(ab[dom]), "pairs_processed": r["P_exam"], "pairs_listed": r["P"]},
    }
    if r.get("block_pairs") is not None:
        line["config"]["block_pairs"] = r["block_pairs"]  # (Gaussian, 64-px block) pairs, what is sorted
    if "ms_fwd_only" in r:
        line["fwd_only_mpix_s"] = round(r["n_pix"] / (r["ms_fwd_only"] / 1e3) / 1e6, 3)
    if not args.no_traffic:
        try:
            tr = measure_traffic(args.workload, args.exact, dom, bwd_opts)
            line["roofline"]["traffic"] = int(tr["bytes"])
            line["roofline"]["traffic_detail"] = {k: (round(v, 3) if isinstance(v, float) else v)
                                                  for k, v in tr.items()}
            # the DRAM bytes the counters measured, over the same launch time
            line["roofline"]["dram_achieved"] = round(tr["bytes"] / t_dom / 1e9, 2)
            line["roofline"]["dram_frac"] = round(tr["bytes"] / t_dom / HBM_PEAK, 4)
        except Exception as e:  # counters unavailable: traffic stays null, say why
            line["roofline"]["traffic_error"] = str(e)
        print(f"[bench] traffic: {line['roofline'].get('traffic')}", file=sys.stderr, flush=True)
        # the compositing kernels are VALU-issue bound, not HBM bound: report that rate and the FP32 FLOP roof
        try:
            line["roofline"]["valu_issue"] = measure_valu(args.workload, args.exact, dom, line["kernels_ms"][dom],
                                                          bwd_opts)
        except Exception as e:
            line["roofline"]["valu_issue"] = {"error": str(e)}
        print(f"[bench] valu: {line['roofline']['valu_issue']}", file=sys.stderr, flush=True)
        try:
            line["roofline"]["fp32_flops"] = measure_flops(args.workload, args.exact, dom, line["kernels_ms"][dom],
                                                           bwd_opts)
        except Exception as e:
            line["roofline"]["fp32_flops"] = {"error": str(e)}
        print(f"[bench] flops: {line['roofline']['fp32_flops']}", file=sys.stderr, flush=True)
    if args.workload == "dense_3m" and not args.no_extra:
        # the batched 8-view 4K training step and one whole training iteration,
        # measured in the same run (a few steps each)
        mv = make_scene("multiview_6m")
        print(f"[bench] workload {mv.name}: N={mv.N}", file=sys.stderr, flush=True)
        rv = gpu_bench_views(mv, 3, 1)
        line["multiview_6m"] = {"workload": mv.name, "ms_per_step": round(rv["ms_step"], 3),
                                "mpix_s": round(rv["n_pix"] / (rv["ms_step"] / 1e3) / 1e6, 3), "views": rv["views"],
                                ("block_pairs" if rv["shift"] else "pairs"): rv["P"], "kernels_ms": {"composite_fwd": round(rv["ms_composite_fwd"], 3),
                                                                 "composite_bwd": round(rv["ms_composite_bwd"], 3)},
                                "steps": 3, "warmup": 1}
        print(f"[bench] multiview: {line['multiview_6m']}", file=sys.stderr, flush=True)
        del mv
        ts = train_step(5, 2)
        line["train_step_3m"] = ts
        print(f"[bench] training iteration: {ts}", file=sys.stderr, flush=True)
    if args.workload == "dense_3m" and not args.no_serving:
        from gaussian_splatting_3d_amd import loadgen

        sv = loadgen.serving_benchmark(levels=tuple(float(x) for x in args.serve_levels.split(",")),
                                       duration_s=args.serve_duration,
                                       log=lambda m: print(m, file=sys.stderr, flush=True))
        line["serving"] = sv
    if not args.no_cpu_baseline and views == 1:
        cb = cpu_baseline(scene)
        line["cpu_baseline"] = {k: cb[k] for k in ("value", "unit", "cores", "kind", "sample", "cpu_model",
                                                   "step_s")}
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


def train_step(steps, warmup):
    """One whole training iteration of main_sh.py:141-240 at 3M Gaussians,
    1920x1080: render, loss 0.8 L1 + 0.2 SSIM (window 11, the conf/*.yaml
    values), adjoint, the visibility count and gradient statistics of
    adaptive control (sh_renderer.py:215-216, 602-623), and the Adam step of
    the optimizer the reference rebuilds every iteration (fresh moments,
    per-group learning rates of conf/garden_sh.yaml:72-76).  Returns the
    timing record."""
    from gaussian_splatting_3d_amd import densify, rasterize, scenes, training

    scene = scenes.dense_3m()
    dev = torch.device("cuda:0")
    c2w, cam = scene.cameras[0]
    cfg = scenes.RasterConfig()
    params = [t.to(dev).contiguous() for t in (scene.mean, scene.qvec, scene.svec_before_activation,
                                               scene.alpha_before_activation, scene.sh_coeffs)]
    target = scenes.target_image(cam.h, cam.w, 3).to(dev).reshape(cam.h, cam.w, 3)
    adam = training.FusedAdam(list(zip(params, (1e-3, 1e-2, 1e-2, 1e-2, 1e-2))), betas=(0.9, 0.99), fresh=True)
    dc = densify.DensityControl(scene.N, device=dev)

    def step():
        out, st = rasterize.forward_view(*params, c2w, cam, scene.C, cfg, prepare_backward=True)
        loss, grad_out = training.photometric_loss(out.view(cam.h, cam.w, 3), target, "l1", ssim_mult=0.2,
                                                   ssim_win=11)
        g = rasterize.backward_view(st, *params, grad_out.view(-1))
        dc.count_visible(st.b.vis)
        dc.update_grads(g[0])
        adam.step(list(g[:5]))
        return loss, st

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        loss, st = step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / steps
    assert math.isfinite(float(loss))
    return {"workload": "train_step_3000k_1920x1080_sh3", "ms_per_step": round(ms, 4), "it_s": round(1e3 / ms, 3),
            "mpix_s": round(cam.h * cam.w / (ms / 1e3) / 1e6, 3), "gaussians": scene.N, "visible": st.b.V,
            ("block_pairs" if st.b.shift else "pairs"): st.b.P, "width": cam.w, "height": cam.h,
            "sh_degree": scene.C - 1, "ssim_mult": 0.2, "ssim_win": 11, "steps": steps, "warmup": warmup}


def train_main(args):
    """The training-iteration workload as its own line."""
    r = train_step(args.steps, args.warmup)
    line = {
        "metric": "training iterations/s @3M Gaussians 1080p (render, L1+SSIM loss, adjoint, density stats, Adam)",
        "value": r["it_s"],
        "unit": "it/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded scene and target, random-init Gaussians)",
        "config": {k: r[k] for k in r if k not in ("ms_per_step", "it_s", "mpix_s", "steps", "warmup")},
        "mpix_s": r["mpix_s"],
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def serve_main(args):
    """Served views/s and p50/p99 latency of the render server against the
    plain per-call renderer on the same seeded request stream."""
    from gaussian_splatting_3d_amd import loadgen

    sv = loadgen.serving_benchmark(levels=tuple(float(x) for x in args.serve_levels.split(",")),
                                   duration_s=args.serve_duration, log=lambda m: print(m, file=sys.stderr, flush=True))
    line = {
        "metric": "served views/s at the plain renderer's p99 (1M+4M SH3 scenes, mixed 480p-1080p)",
        "value": sv["served_views_s_at_plain_p99"],
        "unit": "views/s",
        "n_gpus": 1,
        "steps": None,
        "warmup": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded scenes, Poisson arrivals, random orbit cameras)",
        "config": {"workload": "serve_1m_4m_mixed_res", **sv["config"]},
        "serving": sv,
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)

if __name__ == "__main__":
    main()
