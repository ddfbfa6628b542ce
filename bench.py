"""Benchmark: training throughput of the Factor Fields field path on one MI355X.

Default workload (N=1): NeRF-style per-scene training on the synthetic
Gaussian scene, 4,096 rays per step, in two forms trained the same number of
steps: the learned proposal with BF16 tables and the matrix-core decoder (the
line's value, rays/s) and the plain FP32 form with the occupancy sampler
(timed beside it).  One step = batch gather + sampler + field forward +
compositing + loss + backward + fused Adam.  The other workloads (3D SDF
regression, 2D image regression, dictionary image set) run after it at
bounded sizes and are nested in the same line under config.other_workloads;
--workload sdf|image|dict runs one of them alone as the line's value.

Every leg reports its dominant kernel's roofline fraction and a CPU baseline
(the oracle, the reference's ATen ops on the host cores, on a bounded sample
of the same workload); the SDF and dictionary legs also time their forward
gather kernel alone (forward_gather).  SDF batches are sorted draws from a
Morton-ordered pool by default (--sdf-order random: draw order).  Prints one
JSON line (rank 0).  Only one GPU is used; --gpus N > 1 exits.
"""

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "factor-fields_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = ("train rays/s & field pts/s, proposal+BF16 vs plain FP32 on 1 GPU at ≤0.1 dB PSNR gap")
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def copy_calibration():
    """Three device-to-device copies of 1 GiB each: the reference dispatch for
    scaling rocprofv3's FETCH_SIZE / WRITE_SIZE counters to HBM bytes (each
    copy reads and writes exactly 2^30 bytes)."""
    import torch
    src = torch.empty(1 << 28, device="cuda")
    dst = torch.empty_like(src)
    src.fill_(1.0)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` measured by the two rocprofv3 --pmc
    runs recorded under profiles/ (null when no record exists)."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        rec = json.load(f)
    ent = rec.get("kernels", {}).get(kernel)
    return None if ent is None else ent.get("hbm_bytes_per_launch")


def sdf_bytes_per_point(model):
    """Algorithmic HBM bytes per point of the fused SDF training kernel:
    coordinates (12 B) + target (4 B) + one read of every gathered table value
    (4 B each) + read-modify-write of its FP32 gradient (8 B each)."""
    dims = model.basis_dims
    G = sum(8 * d for d in dims)                       # trilinear basis corners
    C = sum(dims)
    G += 3 * 4 * C + 3 * 2 * C                         # VM: 3 planes x 4 corners + 3 lines x 2
    return 12 + 4 + 4 * G + 8 * G, G


def run_sdf(args):
    import torch
    from ffields import FactorFields, FusedAdam
    from ffields import workloads as W

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    cfg = W.sdf_cfg()
    n = int(cfg.training.batch_size)
    log("building model and data")
    model = FactorFields(cfg, dev)
    scene = W.SdfScene()
    pool_n = 4 * n
    pts, sdf = scene.training_pool(pool_n)
    pts, sdf = pts.to(dev), sdf.to(dev)
    if args.sdf_order == "morton":
        perm = W.morton_order(pts, cfg.dataset.aabb)
        pts, sdf = pts[perm], sdf[perm]
    opt = FusedAdam(model.get_optparam_groups(cfg.training.lr_small, cfg.training.lr_large),
                    betas=(0.9, 0.99))
    gen = torch.Generator(device=dev).manual_seed(W.BASE_SEED + 1)
    loss_sum = torch.zeros((), device=dev)
    x_last = [None]

    def step(ev=None):
        if args.sdf_order == "morton":
            idx = W.sorted_batch(pool_n, n, gen, dev)
        else:
            idx = torch.randint(0, pool_n, (n,), device=dev, generator=gen)
        x, t = pts[idx], sdf[idx]
        x_last[0] = x
        loss_sum.zero_()
        if ev is not None:
            ev[0].record()
        model.regression_step(x, t, loss="l1", loss_sum=loss_sum)
        if ev is not None:
            ev[1].record()
        opt.step(zero_grad=True)

    log(f"warmup {args.warmup}")
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(args.steps)]
    log(f"timing {args.steps} steps")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        step(evs[k])
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    dt = t1 - t0
    kern_ms = sum(a.elapsed_time(b) for a, b in evs) / args.steps
    final_l1 = loss_sum.item() / n
    bpp, G = sdf_bytes_per_point(model)
    achieved = bpp * n / (kern_ms * 1e-3) / 1e9
    # the forward gather kernel alone (field_fwd_kernel: SDF evaluation,
    # FactorFields.forward without autograd) on one batch of the same order
    fx = x_last[0]
    with torch.no_grad():
        for _ in range(3):
            model.field(fx)
        torch.cuda.synchronize()
        f0, f1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        f0.record()
        for _ in range(20):
            model.field(fx)
        f1.record()
        torch.cuda.synchronize()
    fwd_ms = f0.elapsed_time(f1) / 20
    fwd_bpp = 12 + 4 * G + 4                            # coordinates, gathered values, output
    fwd_gbs = fwd_bpp * n / (fwd_ms * 1e-3) / 1e9
    result = {
        "metric": METRIC,
        "value": args.steps * n / dt,
        "unit": "field pts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded sphere/box union SDF)",
        "config": {"workload": "sdf_regression_vm: 2^20 pts/step, 3-D grid basis [4,4,4,2,2,2]@[32..128] "
                               "sawtooth, VM coeff 3x128^2+3x128 (18 ch), MLP 18-64-1, L1, Adam",
                   "batch_points": n, "batch_order": args.sdf_order, "params": model.n_parameters(), "final_l1": final_l1},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic("field_bwd_mfma_kernel (sdf)"),
                     "kernel": "field_bwd_mfma_kernel (fused forward+loss+backward)",
                     "kernel_ms": kern_ms, "bytes_per_point": bpp, "gathered_values_per_point": G},
        "forward_gather": {"kernel": "field_fwd_kernel (forward only, one batch of the same order)",
                           "bound": "hbm", "kernel_ms": fwd_ms, "bytes_per_point": fwd_bpp, "achieved": fwd_gbs,
                           "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": fwd_gbs / HBM_PEAK_GBS},
    }
    return result, model


def run_atomics(args):
    """The FP32 atomic line-request rate (ff_atomic_rate_probe): 2^26 line
    requests into a 64 MiB target at 4, 16 and 64 lines per wave
    instruction, timed with HIP events on the launch stream.  The ceiling the
    backward kernels' table scatters are compared with."""
    import ctypes
    import torch
    import ffields._lib as L
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    n_lines = 1 << 20
    dst = torch.zeros(n_lines * 16, device=dev)
    total = 1 << 26
    rates = {}
    st = torch.cuda.current_stream()
    for lpi in (4, 16, 64):
        n_instr = total // lpi
        for _ in range(max(args.warmup, 1)):
            L.check(L.lib().ff_atomic_rate_probe(L.ptr(dst), n_lines, n_instr, lpi, L.stream_of(dst)), "probe")
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(st)
        for _ in range(args.steps):
            L.check(L.lib().ff_atomic_rate_probe(L.ptr(dst), n_lines, n_instr, lpi, L.stream_of(dst)), "probe")
        b.record(st)
        torch.cuda.synchronize()
        ms = a.elapsed_time(b) / args.steps
        rates[lpi] = {"ms_per_launch": ms, "g_line_requests_per_s": total / (ms * 1e-3) / 1e9}
        log(f"atomics: {lpi} lines/instruction: {rates[lpi]['g_line_requests_per_s']:.2f} G line requests/s")
    best = max(r["g_line_requests_per_s"] for r in rates.values())
    return {"metric": METRIC, "value": best, "unit": "G atomic line requests/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": rates[64]["ms_per_launch"], "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic (hashed line indices)",
            "config": {"workload": "fp32 atomic line-request rate: 2^26 requests into 64 MiB", "rates": rates},
            "roofline": {"bound": "atomics", "achieved": best, "peak": None, "unit": "G line requests/s",
                         "frac": None, "traffic": None, "kernel": "atomic_rate_kernel"},
            "cpu_baseline": None}, None


def run_mesh(args):
    """Inference-only kernels of the export and occupancy paths.

    Mesh export (utils.py:183-249 via train_per_scene.py:32-41): marching
    tetrahedra of a 256^3 FP32 signed-distance volume of the synthetic SDF
    scene (32 spheres and 16 boxes), the whole ffields.mesh.marching_tetrahedra
    call (count, scan, emit, key sort, vertices) timed with HIP events; the
    count kernel (ff_mt_count, 8 B per cube: one 4 B value read per cube, its
    neighbours' from cache, and the 4 B count written) timed alone for the
    roofline.  Occupancy filter (models/FactorFields.py:774-775): the
    small-object filter ff_remove_small_components on the default 200^3 grid
    of a ball, a shell and 2 % scattered floaters.  CPU baselines: the
    oracle's marching tetrahedra on a 96^3 crop of the same volume and the
    oracle's scipy labelling of the same mask."""
    import ctypes
    import numpy as np
    import torch
    import ffields._lib as L
    from ffields import workloads as W
    from ffields.mesh import marching_tetrahedra
    from ffields.render import remove_small_components

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    n = 256
    scene = W.SdfScene()
    g = torch.linspace(-1, 1, n, dtype=torch.float64)
    vol = torch.empty((n, n, n))
    for i in range(n):                       # slabs: the scene's SDF is evaluated on the host
        y, z = torch.meshgrid(g, g, indexing="ij")
        pts = torch.stack([torch.full_like(y, float(g[i])), y, z], -1).reshape(-1, 3)
        vol[i] = scene.sdf(pts).reshape(n, n).float()
    vol = vol.to(dev)

    def timed(fn, reps):
        for _ in range(max(args.warmup, 1)):
            fn()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(reps):
            fn()
        b.record()
        torch.cuda.synchronize()
        return a.elapsed_time(b) / reps

    verts, faces = marching_tetrahedra(vol, 0.0, inside_above=False)
    mesh_ms = timed(lambda: marching_tetrahedra(vol, 0.0, inside_above=False), args.steps)
    cubes = (n - 1) ** 3
    counts = torch.empty(cubes, dtype=torch.int32, device=dev)
    lev = ctypes.c_double(0.0)
    count_ms = timed(lambda: L.check(L.lib().ff_mt_count(L.ptr(vol), n, n, n, lev, 0, L.ptr(counts),
                                                          L.stream_of(vol)), "ff_mt_count"), args.steps)
    log(f"mesh: {n}^3, {faces.shape[0]} triangles, {verts.shape[0]} vertices, {mesh_ms:.3f} ms "
        f"(count kernel {count_ms * 1e3:.1f} us)")
    count_gbs = 8 * cubes / (count_ms * 1e-3) / 1e9

    m = 200
    gm = np.linspace(-1, 1, m)
    zz, yy, xx = np.meshgrid(gm, gm, gm, indexing="ij")
    r = np.sqrt(xx * xx + yy * yy + zz * zz)
    mask_np = (r < 0.3) | ((r > 0.6) & (r < 0.62)) | (np.random.default_rng(3).random((m, m, m)) < 0.02)
    mask = torch.from_numpy(mask_np.astype(np.uint8)).to(dev)
    keep = remove_small_components(mask, float(m))
    cc_ms = timed(lambda: remove_small_components(mask, float(m)), args.steps)
    log(f"components: {m}^3, {int(keep.sum())} of {int(mask.sum())} voxels kept, {cc_ms:.3f} ms")

    out = {"metric": METRIC, "value": cubes / (mesh_ms * 1e-3), "unit": "grid cubes/s (mesh extraction)",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": mesh_ms,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32/f64",
           "data": "synthetic (seeded sphere/box union SDF on a 256^3 grid; ball + shell + 2% floaters mask)",
           "config": {"workload": "mesh export: marching tetrahedra of a 256^3 SDF volume; occupancy small-object "
                                  "filter on 200^3",
                      "triangles": int(faces.shape[0]), "vertices": int(verts.shape[0]),
                      "components": {"ms": cc_ms, "voxels_per_s": m ** 3 / (cc_ms * 1e-3),
                                     "kept": int(keep.sum()), "set": int(mask.sum())}},
           "roofline": {"bound": "hbm", "achieved": count_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": count_gbs / HBM_PEAK_GBS, "traffic": None, "kernel": "mt_count_kernel",
                        "kernel_ms": count_ms, "bytes_per_cube": 8},
           "cpu_baseline": None}
    if not args.no_cpu_baseline:
        from oracle.mesh import marching_tetrahedra as oracle_mt
        from oracle.render import remove_small_objects
        torch.set_num_threads(min(16, os.cpu_count() or 1))
        crop = vol[:96, :96, :96].double().cpu()
        t0 = time.perf_counter()
        oracle_mt(crop, 0.0, inside_above=False)
        mt_s = time.perf_counter() - t0
        t0 = time.perf_counter()
        remove_small_objects(mask_np, float(m))
        cc_s = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": 95 ** 3 / mt_s, "unit": "grid cubes/s (mesh extraction)",
                               "cores": torch.get_num_threads(), "kind": "port",
                               "components_voxels_per_s": m ** 3 / cc_s,
                               "sample": "oracle marching tetrahedra (torch CPU) on a 96^3 crop of the same volume; "
                                         "oracle small-object filter (scipy.ndimage.label, one thread) on the same "
                                         "200^3 mask"}
    return out


def nerf_setup(dev, pool_rays=1 << 20, seed_off=0):
    import torch
    from ffields import workloads as W
    scene = W.NerfScene(device=dev)
    g = torch.Generator(device=dev).manual_seed(W.BASE_SEED + 2000 + seed_off)
    rays = scene.random_rays(pool_rays, g)
    rgb = scene.render_gt(rays)
    return scene, rays, rgb


def nerf_bytes_per_sample(model, table_bytes):
    """Algorithmic HBM bytes per evaluated sample of the field backward, the
    per-unit model of the workload definition: position (4 B x 3) + the
    sample's upstream gradient (sigma, rgb: 4 B x 4) + one read of every
    gathered table value (`table_bytes` each) + read-modify-write of its FP32
    gradient (8 B each), with G = 8 corners per basis channel + 8 corners per
    coefficient channel = 288.  FP32: 3,484 B; BF16: 2,908 B.  (The kernels
    read the values the forward stored instead of gathering again, so they
    move fewer bytes than this count credits.)"""
    C = sum(model.basis_dims)
    G = 8 * C + 8 * C
    return 12 + 16 + table_bytes * G + 8 * G, G


def nerf_train(form, dev, pool, pool_rgb, test_rays, test_rgb, timed_steps, pretrain, warmup, check_rays=256,
               heldout=None):
    """Train one NeRF form for `pretrain` steps (building the occupancy grid
    halfway), run `warmup` untimed steps, then time `timed_steps` steps.

    form 'plain': FP32 tables, uniform sampler with the 128^3 occupancy grid
    (train_per_scene.py:126-200 step order: sample, field, composite, MSE,
    backward, Adam).  form 'proposal': BF16 table copies read by the
    kernels (FP32 masters updated by Adam), the learned proposal placing 48
    fine samples per ray and trained on the field's rendering weights each
    step; once the occupancy grid exists, the occupancy sampler renders the
    first `check_rays` rays of every batch (side stream), their error is
    compared with the proposal's training render of the same rays, and the
    proposal is switched off (training falls back to the occupancy sampler)
    if the accumulated PSNR gap exceeds 0.1 dB."""
    import torch
    from ffields import FactorFields, FusedAdam
    from ffields import workloads as W
    from ffields.params import N_to_reso, cal_n_samples
    from ffields.proposal import ProposalReferee
    from ffields.rays import gather_batch

    cfg = W.nerf_cfg()
    torch.manual_seed(W.BASE_SEED)
    prop = form == "proposal"
    model = FactorFields(cfg, dev, table_dtype="bf16" if prop else "fp32")
    B = int(cfg.training.batch_size)
    reso = N_to_reso(cfg.training.volume_resoInit ** 3, model.aabb.cpu())
    n_samples = min(int(cfg.renderer.max_samples), cal_n_samples(reso, cfg.renderer.step_ratio))
    if prop:
        model.enable_proposal(reso=64, n_fine=48, table_dtype="bf16")
    opt = FusedAdam(model.get_optparam_groups(cfg.training.lr_small, cfg.training.lr_large), betas=(0.9, 0.99),
                    shadow_of=model.shadow_lookup)
    lr_factor = 0.1 ** (1 / cfg.training.n_iters)
    gen = torch.Generator(device=dev).manual_seed(W.BASE_SEED + 2001)
    counts = torch.zeros(max(timed_steps, 1), dtype=torch.int32, device=dev)
    referee = ProposalReferee(dev, n_rays=check_rays, max_gap_db=0.1, n_samples=n_samples) if prop else None
    state = {"fallback_at": None}

    def proposal_on():
        return referee is not None and referee.active

    def step(k=None):
        idx = torch.randint(0, pool.shape[0], (B,), device=dev, generator=gen)
        rays, rgb = gather_batch(pool, pool_rgb, idx)
        if proposal_on():
            referee.compare(model, rays, rgb)          # side stream, overlaps the training pass
            model.render_train_step_proposal(rays, rgb, white_bg=True, reduce=False)
            referee.add_proposal(model.last_rgb_map(), rgb)
            referee.join()
        else:
            model.render_train_step(rays, rgb, white_bg=True, N_samples=n_samples, reduce=False)
        if k is not None:
            counts[k] = model._rbuf[1]["ray_off"][-1]
        opt.step(zero_grad=True)
        for gp in opt.param_groups:
            gp["lr"] *= lr_factor

    def referee_check(it):
        if proposal_on() and referee.decide() is not None and not referee.active:
            state["fallback_at"] = it

    upd = pretrain // 2
    for it in range(pretrain):
        step()
        if it == upd:
            model.updateAlphaMask(tuple(reso), is_update_alphaMask=True)
        if it > upd and (it - upd) % 100 == 0:
            referee_check(it)
    referee_check(pretrain)
    occ_frac = float(model.alphaMask.occ.float().mean()) if model.alphaMask is not None else 1.0
    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    model.kernel_events = []
    t0 = time.perf_counter()
    for k in range(timed_steps):
        step(k)
    t_enq = time.perf_counter() - t0
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    kern = model.kernel_events
    model.kernel_events = None
    kern_ms = sum(a.elapsed_time(b) for a, b in kern) / max(len(kern), 1)
    samples = int(counts[:timed_steps].sum())
    with torch.no_grad():
        out = [model.render_eval(test_rays[k:k + 4096], use_proposal=proposal_on(), N_samples=n_samples)
               for k in range(0, test_rays.shape[0], 4096)]
        mse = torch.mean((torch.cat(out) - test_rgb) ** 2).item()
        heldout_psnr = []
        for hr, hg in (heldout or []):
            img = torch.cat([model.render_eval(hr[k:k + 16384], use_proposal=proposal_on(), N_samples=n_samples)
                             for k in range(0, hr.shape[0], 16384)])
            heldout_psnr.append(-10 * math.log10(torch.mean((img - hg) ** 2).item()))
    bps, G = nerf_bytes_per_sample(model, 2 if prop else 4)
    samples_per_launch = samples / max(timed_steps, 1)
    res = {"rays_per_s": timed_steps * B / dt, "field_pts_per_s": samples / dt,
           "ms_per_step": dt / timed_steps * 1e3, "host_enqueue_ms_per_step": t_enq / timed_steps * 1e3,
           "samples_per_ray": samples / (timed_steps * B),
           "ns_per_sample": dt / max(samples, 1) * 1e9,
           "test_psnr": -10 * math.log10(mse), "proposal_active": proposal_on(),
           "fallback_at_step": state["fallback_at"],
           "sampler_gaps_db": referee.gaps_db[-5:] if referee is not None else [],
           "sampler_gap_db_stats": ({"checks": len(referee.gaps_db), "min": min(referee.gaps_db),
                                     "max": max(referee.gaps_db),
                                     "mean": sum(referee.gaps_db) / len(referee.gaps_db)}
                                    if referee is not None and referee.gaps_db else None),
           "train_steps": pretrain + warmup + timed_steps, "occupancy_fraction": occ_frac, "n_samples": n_samples,
           "table_dtype": "bf16" if prop else "fp32", "params": model.n_parameters(),
           "heldout_views_psnr": heldout_psnr or None,
           "heldout_psnr_mean": sum(heldout_psnr) / len(heldout_psnr) if heldout_psnr else None,
           "field_bwd_ms": kern_ms, "field_bwd_bytes_per_sample": bps, "gathered_values_per_sample": G,
           "field_bwd_gbs": bps * samples_per_launch / (kern_ms * 1e-3) / 1e9 if kern_ms else None}
    return res, model


def run_nerf(args):
    """NeRF-style per-scene training on the synthetic Gaussian scene: the
    proposal+BF16 form is the measured value, the plain FP32 form (same
    scene, model, batch and step count) is timed beside it."""
    import torch

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    log("nerf: ground-truth ray pool")
    _, pool, pool_rgb = nerf_setup(dev, pool_rays=args.nerf_pool)
    _, test_rays, test_rgb = nerf_setup(dev, pool_rays=1 << 14, seed_off=7)
    heldout = None
    if args.nerf_test_views > 0:
        # whole images from cameras that are not among the 100 training views
        from ffields import workloads as W
        res_px = args.nerf_test_res
        hs = W.NerfScene(n_views=args.nerf_test_views, H=res_px, W=res_px, device=dev)
        log(f"nerf: {args.nerf_test_views} held-out views at {res_px}x{res_px}")
        heldout = []
        for v in range(args.nerf_test_views):
            r = hs.rays(v, torch.arange(res_px * res_px, device=dev))
            heldout.append((r, hs.render_gt(r)))
    forms = {"both": ["plain", "proposal"], "plain": ["plain"], "proposal": ["proposal"]}[args.nerf_form]
    res, models = {}, {}
    for f in forms:
        log(f"nerf {f}: {args.nerf_pretrain} training steps, {args.warmup} warmup, then {args.steps} timed")
        res[f], models[f] = nerf_train(f, dev, pool, pool_rgb, test_rays, test_rgb, args.steps, args.nerf_pretrain,
                                       args.warmup, heldout=heldout)
        log(f"nerf {f}: {res[f]['rays_per_s']:.0f} rays/s, PSNR {res[f]['test_psnr']:.2f} dB")
    main = res[forms[-1]]
    out = {
        "metric": METRIC,
        "value": main["rays_per_s"],
        "unit": "rays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": main["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        # proposal form: BF16 table gathers and BF16 matrix-core decoder with
        # FP32 accumulation; plain form: FP32 throughout
        "dtype": "bf16" if forms[-1] == "proposal" else "f32",
        "data": "synthetic (48 seeded anisotropic Gaussians, sinusoid colour; ground truth by 1,024-sample "
                "quadrature on random rays from 100 radius-4 cameras)",
        "config": {"workload": ("nerf_proposal_bf16" if forms[-1] == "proposal" else "nerf_plain_fp32")
                   + ": 4,096 rays/batch, 3-D grid basis [4,4,4,2,2,2]@[32..128], 32^3x18 coefficient grid, "
                     "18-64-28 decoder + degree-2 SH head",
                   "field_pts_per_s": main["field_pts_per_s"], "pretrain_steps": args.nerf_pretrain,
                   "forms": res},
    }
    if "proposal" in res and "plain" in res:
        pr, pl = res["proposal"], res["plain"]
        out["config"]["speedup_vs_plain_fp32"] = pr["rays_per_s"] / pl["rays_per_s"]
        # rays/s ratio = (samples per ray, plain / proposal) x (time per sample, plain / proposal)
        out["config"]["speedup_split"] = {"samples_per_ray_ratio": pl["samples_per_ray"] / pr["samples_per_ray"],
                                          "per_sample_time_ratio": pl["ns_per_sample"] / pr["ns_per_sample"]}
        out["config"]["psnr_gap_db"] = pl["test_psnr"] - pr["test_psnr"]
    bw = main["field_bwd_gbs"]
    # the plain form's backward is two kernels (decoder on the stored rows,
    # then the table scatter); its traffic is the sum of their records
    knames = (["nerf_field_bwd_mfma_kernel"] if forms[-1] == "proposal"
              else ["nerf_decoder_bwd_kernel", "nerf_scatter_kernel"])
    tr = [measured_traffic(k) for k in knames]
    out["roofline"] = {"bound": "hbm", "achieved": bw, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": bw / HBM_PEAK_GBS if bw else None,
                       "traffic": None if None in tr else sum(tr),
                       "kernel": " + ".join(knames) + " (+ the decoder-gradient reduction)",
                       "kernel_ms": main["field_bwd_ms"], "bytes_per_sample": main["field_bwd_bytes_per_sample"]}
    return out, models.get("plain")


def run_dict(args):
    """Dictionary image set: 1,024 seeded 256x256 images sharing one 144-channel
    2-D basis with per-image 32x32x144 coefficient grids, decoder 144-64-3,
    2^22 pixels per step in one launch, MSE, Adam.  Pixels and images are
    drawn uniformly with replacement (dataLoader/image_set.py:55-56); with
    --dict-order cell (default) each batch arrives sorted by image and
    coefficient cell (same distribution, drawn inside the timed step), so the
    fused kernel merges the atomics of pixels that share texels.  The BF16
    form is the line's value; the FP32 form (FP32 tables, FP32 matrix-core
    decoder, the same fused kernel layout) is trained on the same batches and
    timed beside it.

    Accuracy: at the reference's constant learning rate the PSNR of this fit
    swings by several dB within a few hundred steps, so one evaluation says
    little.  Both forms are scored on 16 whole images every --dict-eval-every
    steps over the second half of training and compared by the mean of those
    scores; the FP32 form's scores are also taken through the BF16 kernels
    (its FP32 tables and weights loaded into a BF16-form model), which is the
    BF16 form's forward precision cost on the same parameters; with
    --dict-seeds 2 a second FP32 run on other batches gives the FP32 form's
    own run-to-run spread of that mean."""
    import torch
    from ffields import FactorFields, FusedAdam
    from ffields import workloads as W

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    N, S = args.dict_images, 256
    log(f"dict: generating {N} images")
    imgs = W.synth_image_set(N, S, dev)
    cfg = W.dict_cfg(N, S)
    n = args.dict_batch
    eval_imgs = torch.arange(0, N, max(1, N // 16), device=dev)[:16].tolist()
    pix = torch.arange(S * S, device=dev)
    ex = torch.stack([(pix % S).float() + 0.5, (pix // S).float() + 0.5], -1)
    eval_x = [torch.cat([ex, torch.full((S * S, 1), i + 0.5, device=dev)], -1) for i in eval_imgs]

    order = W.cell_order(S, int(cfg.model.coeff_reso), dev) if args.dict_order == "cell" else None

    def batch(gen):
        img, p = W.image_set_batch(n, N, S, gen, dev, order=order)
        x = torch.stack([(p % S).float() + 0.5, (p // S).float() + 0.5, img.float() + 0.5], -1)
        return x, imgs[img, p]

    @torch.no_grad()
    def psnr_of(model):
        se = 0.0
        for x, i in zip(eval_x, eval_imgs):
            se += float(((model.field(x) - imgs[i]) ** 2).sum())
        return -10 * math.log10(se / (len(eval_imgs) * S * S * 3))

    seeds = [W.BASE_SEED + 4001, W.BASE_SEED + 9001]
    runs = [("bf16", seeds[0])]
    if args.dict_fp32:
        runs.append(("fp32", seeds[0]))
        if args.dict_seeds > 1:
            runs.append(("fp32", seeds[1]))
    res, spread = {}, None
    for form, seed in runs:
        torch.manual_seed(W.BASE_SEED)
        model = FactorFields(cfg, dev, table_dtype=form)
        opt = FusedAdam(model.get_optparam_groups(cfg.training.lr_small, cfg.training.lr_large),
                        betas=(0.9, 0.99), shadow_of=model.shadow_of)
        gen = torch.Generator(device=dev).manual_seed(seed)
        fused = model.fused_supported()
        primary = seed == seeds[0]
        # FP32 parameters through the BF16 kernels (forward precision cost)
        probe = FactorFields(cfg, dev, table_dtype="bf16") if form == "fp32" and primary else None

        def step(ev=None):
            x, t = batch(gen)
            if ev is not None:
                ev[0].record()
            if fused:
                model.regression_step(x, t, loss="mse", reduce=False)
            else:
                ((model.field(x) - t) ** 2).mean().backward()
            if ev is not None:
                ev[1].record()
            opt.step(zero_grad=True)

        log(f"dict {form} (batch seed {seed}): {args.dict_steps} training steps "
            f"({'fused' if fused else 'generic'} kernels)")
        scores, costs = [], []
        t_tr = time.perf_counter()
        for it in range(args.dict_steps):
            step()
            s_ = it + 1
            if s_ > args.dict_steps // 2 and (s_ % args.dict_eval_every == 0 or s_ == args.dict_steps):
                scores.append(psnr_of(model))
                if probe is not None:
                    probe.load_state_dict(model.state_dict())
                    costs.append(scores[-1] - psnr_of(probe))
            if s_ % 100 == 0:
                torch.cuda.synchronize()
                log(f"dict {form}: step {s_}, {(time.perf_counter() - t_tr) / s_ * 1e3:.1f} ms/step"
                    + (f", late PSNR mean {sum(scores) / len(scores):.2f} dB" if scores else ""))
        torch.cuda.synchronize()
        if not primary:
            spread = sum(scores) / len(scores) - res["fp32"]["psnr_late_mean"]
            log(f"dict fp32, other batches: late PSNR mean {sum(scores) / len(scores):.3f} dB")
            del model, opt
            torch.cuda.empty_cache()
            continue
        timed = args.steps if (form == "bf16" or fused) else 0
        evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(timed)]
        t0 = time.perf_counter()
        for k in range(timed):
            step(evs[k])
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        r = {"train_steps": args.dict_steps + timed, "psnr_16_images": psnr_of(model),
             "psnr_late_mean": sum(scores) / len(scores), "psnr_late_evals": len(scores),
             "psnr_late_min": min(scores), "psnr_late_max": max(scores),
             "kernels": "fused" if fused else "generic", "params": model.n_parameters()}
        if costs:
            r["bf16_forward_cost_db"] = sum(costs) / len(costs)
        if timed:
            r["pts_per_s"] = timed * n / dt
            r["ms_per_step"] = dt / timed * 1e3
            r["train_kernel_ms"] = sum(a.elapsed_time(b) for a, b in evs) / timed
        if timed and form == "bf16":
            # the forward gather kernel alone on one batch of the same order
            fx, _ = batch(gen)
            with torch.no_grad():
                for _ in range(3):
                    model.field(fx)
                torch.cuda.synchronize()
                f0, f1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                f0.record()
                for _ in range(10):
                    model.field(fx)
                f1.record()
                torch.cuda.synchronize()
            r["forward_kernel_ms"] = f0.elapsed_time(f1) / 10
        res[form] = r
        log(f"dict {form}: late PSNR mean {r['psnr_late_mean']:.3f} dB over {len(scores)} evaluations"
            + (f", {r['pts_per_s']:.3e} pts/s" if timed else ""))
        del model, opt, probe
        torch.cuda.empty_cache()
    b = res["bf16"]
    bpp = 12 + 12 + 2 * 1152 + 8 * 1152
    achieved = bpp * n / (b["train_kernel_ms"] * 1e-3) / 1e9
    out = {"metric": METRIC, "value": b["pts_per_s"], "unit": "field pts/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.dict_steps, "ms_per_step": b["ms_per_step"], "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "bf16",
           "data": f"synthetic ({N} seeded Gaussian+sinusoid images, 256x256)",
           "config": {"workload": f"dictionary_image_set_bf16: {N} images, basis [32,32,32,16,16,16]@[32..128], "
                                  "32x32x144 coefficient per image, decoder 144-64-3, 2^22 px/step",
                      "batch_order": args.dict_order,
                      "forms": res},
           "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic("dict_bwd_kernel"),
                        "kernel": "dict_bwd_kernel (+ its decoder-gradient reduction)",
                        "kernel_ms": b["train_kernel_ms"], "bytes_per_point": bpp}}
    if "forward_kernel_ms" in b:
        fwd_bpp = 12 + 2 * 1152 + 12                    # coordinates, gathered BF16 values, rgb
        fwd_gbs = fwd_bpp * n / (b["forward_kernel_ms"] * 1e-3) / 1e9
        out["forward_gather"] = {"kernel": "dict_fwd_kernel_bf16 (forward only, one batch of the same order)",
                                 "bound": "hbm", "kernel_ms": b["forward_kernel_ms"], "bytes_per_point": fwd_bpp,
                                 "achieved": fwd_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                 "frac": fwd_gbs / HBM_PEAK_GBS}
    if "fp32" in res:
        out["config"]["psnr_gap_db"] = res["fp32"]["psnr_late_mean"] - b["psnr_late_mean"]
        out["config"]["bf16_forward_cost_db"] = res["fp32"].get("bf16_forward_cost_db")
        if spread is not None:
            out["config"]["fp32_run_spread_db"] = spread
    return out


def image_bytes_per_point(model, out_dim=3):
    """Algorithmic HBM bytes per pixel of the fused 2-D regression kernel:
    coordinates (8 B) + target (4 B per channel) + one read of every
    gathered table value (4 B each) + read-modify-write of its FP32 gradient
    (8 B each); G = 4 bilinear corners per basis and coefficient channel."""
    C = sum(model.basis_dims)
    G = 4 * C + 4 * C
    return 8 + 4 * out_dim + 4 * G + 8 * G, G


def run_image(args):
    """2D image regression: 512x512x3 synthetic target, 65,536 uniformly sampled
    pixels per step, 2-D grid basis [4,4,4,2,2,2]@[32..128] (sawtooth, literal
    freq bands), 64x64x18 coefficient grid, MLP 18-64-3, MSE, Adam lr 0.02,
    2,000 steps (all timed); reports points/s, PSNR over the full image and the
    fused training kernel's roofline."""
    import torch
    from ffields import FactorFields, FusedAdam
    from ffields import workloads as W

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    cfg = W.image_cfg()
    img = W.synth_image(device=dev)
    H, Wd = img.shape[:2]
    coords = W.pixel_coords(H, Wd, device=dev)
    rgb = img.reshape(-1, 3)
    model = FactorFields(cfg, dev)
    opt = FusedAdam(model.get_optparam_groups(cfg.training.lr_small, cfg.training.lr_large), betas=(0.9, 0.99))
    n = int(cfg.training.batch_size)
    gen = torch.Generator(device=dev).manual_seed(W.BASE_SEED + 3)
    iters = args.image_iters

    def step(ev=None):
        idx = torch.randint(0, H * Wd, (n,), device=dev, generator=gen)
        x, t = coords[idx], rgb[idx]
        if ev is not None:
            ev[0].record()
        model.regression_step(x, t, loss="mse", reduce=False)
        if ev is not None:
            ev[1].record()
        opt.step(zero_grad=True)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(iters)]
    t0 = time.perf_counter()
    for k in range(iters):
        step(evs[k])
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    kern_ms = sum(a.elapsed_time(b) for a, b in evs) / iters
    with torch.no_grad():
        out = model.field(coords).clamp(0, 1)
        mse = torch.mean((out - rgb) ** 2).item()
    bpp, G = image_bytes_per_point(model)
    achieved = bpp * n / (kern_ms * 1e-3) / 1e9
    return {
        "metric": METRIC, "value": iters * n / dt, "unit": "field pts/s", "n_gpus": 1, "steps": iters,
        "warmup": args.warmup, "ms_per_step": dt / iters * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded Gaussian+sinusoid image, 512x512)",
        "config": {"workload": "image_regression: 512x512x3, 65,536 px/step, 2-D grid basis [4,4,4,2,2,2]@[32..128] "
                               "sawtooth, 64x64x18 coefficient, MLP 18-64-3, MSE, Adam 0.02",
                   "psnr_full_image": -10 * math.log10(mse), "params": model.n_parameters()},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic("field_bwd_mfma_kernel (image)"),
                     "kernel": "field_bwd_mfma_kernel (fused forward+loss+backward)",
                     "kernel_ms": kern_ms, "bytes_per_point": bpp, "gathered_values_per_point": G},
    }


def _cpu_model():
    """The host CPU's model name (/proc/cpuinfo), for the baseline's record."""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


def cpu_baseline_sdf(args):
    """The oracle (same ATen ops as the reference: grid_sample per level,
    autograd, torch.optim.Adam) on the host cores, bounded sample."""
    import torch
    from ffields import workloads as W
    from oracle import OracleField

    cores = len(os.sched_getaffinity(0))
    torch.set_num_threads(cores)
    cfg = W.sdf_cfg()
    of = OracleField(cfg).requires_grad_(True)
    opt = torch.optim.Adam([{"params": [w for w, _ in of.mlp] + [b for _, b in of.mlp if b is not None],
                             "lr": cfg.training.lr_small},
                            {"params": of.coeffs + of.basises, "lr": cfg.training.lr_large}],
                           betas=(0.9, 0.99))
    scene = W.SdfScene()
    n = args.cpu_batch
    pts, sdf = scene.training_pool(4 * n)
    g = torch.Generator().manual_seed(1)

    def step():
        idx = torch.randint(0, 4 * n, (n,), generator=g)
        out = of.field(pts[idx])
        loss = (out - sdf[idx]).abs().mean()
        opt.zero_grad()
        loss.backward()
        opt.step()

    for _ in range(2):
        step()
    times = []
    t_start = time.perf_counter()
    while len(times) < 5 or (time.perf_counter() - t_start < 10 and len(times) < 50):
        t0 = time.perf_counter()
        step()
        times.append(time.perf_counter() - t0)
        if time.perf_counter() - t_start > 30:
            break
    times.sort()
    med = times[len(times) // 2]
    return {"value": n / med, "unit": "field pts/s", "cores": cores, "kind": "port",
            "sample": f"{len(times)} timed steps (median) of {n} points of the same SDF workload, "
                      f"oracle (torch CPU, grid_sample per level + autograd + Adam), {cores} threads on "
                      f"{_cpu_model()}"}


def _oracle_copy(model):
    """The oracle (CPU, channel-major tables) holding the GPU model's current
    parameters, for a CPU baseline on the same trained state."""
    import copy
    import torch
    from oracle import OracleField
    of = OracleField(copy.deepcopy(model.cfg))
    b, c = model._tables()
    of.basises = [t.detach().float().cpu().contiguous() for t in b]
    of.coeffs = [t.detach().float().cpu().contiguous() for t in c]
    lin = model.linear_mat.backbone
    of.mlp = [(lin[i].weight.detach().cpu().clone(),
               None if lin[i].bias is None else lin[i].bias.detach().cpu().clone()) for i in range(len(lin))]
    return of.requires_grad_(True)


def _oracle_adam(of, cfg):
    import torch
    return torch.optim.Adam([{"params": [w for w, _ in of.mlp] + [b for _, b in of.mlp if b is not None],
                              "lr": cfg.training.lr_small},
                             {"params": of.coeffs + of.basises, "lr": cfg.training.lr_large}],
                            betas=(0.9, 0.99))


def _time_cpu(step, min_steps=3, max_steps=20, budget_s=30.0):
    """Median wall time of `step` after one untimed call: at least min_steps
    timed calls, more while under 10 s, never past budget_s."""
    step()
    times = []
    t_start = time.perf_counter()
    while len(times) < min_steps or (time.perf_counter() - t_start < 10 and len(times) < max_steps):
        t0 = time.perf_counter()
        step()
        times.append(time.perf_counter() - t0)
        if time.perf_counter() - t_start > budget_s:
            break
    times.sort()
    return times[len(times) // 2], len(times)


def cpu_baseline_nerf(args, model):
    """The oracle NeRF path (the reference's ATen ops: sample_point,
    AlphaGridMask grid_sample, grid_sample per level, linear_mat, raw2alpha,
    SH head, autograd, torch.optim.Adam) on the host cores, on the plain GPU
    form's trained state: the same tables, decoder, 128^3 occupancy grid and
    sample count, on a bounded batch of rays of the same scene."""
    import torch
    from ffields import workloads as W
    from ffields.params import N_to_reso, cal_n_samples
    from oracle.render import OracleRenderer

    cores = len(os.sched_getaffinity(0))
    torch.set_num_threads(cores)
    cfg = model.cfg
    of = _oracle_copy(model)
    reso = N_to_reso(cfg.training.volume_resoInit ** 3, of.aabb)
    n_samples = min(int(cfg.renderer.max_samples), cal_n_samples(reso, cfg.renderer.step_ratio))
    am = model.alphaMask
    rd = OracleRenderer(of, None if am is None else am.alpha_volume.cpu()[0, 0],
                        None if am is None else am.aabb.cpu(), torch.tensor(model._step, dtype=torch.float32),
                        n_samples)
    opt = _oracle_adam(of, cfg)
    scene = W.NerfScene(device="cpu")
    g = torch.Generator().manual_seed(W.BASE_SEED + 2000)
    n = args.cpu_rays
    rays = scene.random_rays(n, g)
    rgb = scene.render_gt(rays)
    evaluated = []

    def one():
        out, _, aux = rd.forward(rays, jitter=torch.rand(n), white_bg=True, n_samples=n_samples)
        evaluated.append(int(aux["ray_valid"].sum()))
        loss = torch.mean((out - rgb) ** 2)
        opt.zero_grad()
        loss.backward()
        opt.step()

    med, k = _time_cpu(one)
    spr = sum(evaluated) / (len(evaluated) * n)
    return {"value": n / med, "unit": "rays/s", "cores": cores, "kind": "port",
            "field_pts_per_s": spr * n / med, "samples_per_ray": spr,
            "sample": f"{k} timed steps (median) of {n} rays of the same scene on the plain form's trained "
                      f"state (same tables, decoder and 128^3 occupancy grid, {n_samples} candidates, "
                      f"{spr:.0f} evaluated samples/ray), oracle (torch CPU, autograd + Adam), {cores} threads "
                      f"on {_cpu_model()}"}


def cpu_baseline_image(args):
    """The oracle's 2D image regression step (grid_sample per level, MLP,
    MSE, autograd, Adam) on the host cores, same image, model and batch."""
    import torch
    from ffields import workloads as W
    from oracle import OracleField

    cores = len(os.sched_getaffinity(0))
    torch.set_num_threads(cores)
    cfg = W.image_cfg()
    of = OracleField(cfg).requires_grad_(True)
    opt = _oracle_adam(of, cfg)
    img = W.synth_image()
    coords = W.pixel_coords(img.shape[0], img.shape[1])
    rgb = img.reshape(-1, 3)
    n = int(cfg.training.batch_size)
    g = torch.Generator().manual_seed(W.BASE_SEED + 3)

    def one():
        idx = torch.randint(0, rgb.shape[0], (n,), generator=g)
        loss = torch.mean((of.field(coords[idx]) - rgb[idx]) ** 2)
        opt.zero_grad()
        loss.backward()
        opt.step()

    med, k = _time_cpu(one)
    return {"value": n / med, "unit": "field pts/s", "cores": cores, "kind": "port",
            "sample": f"{k} timed steps (median) of {n} pixels of the same image and model, oracle (torch CPU, "
                      f"grid_sample per level + autograd + Adam), {cores} threads on {_cpu_model()}"}


def cpu_baseline_dict(args):
    """The oracle's dictionary image-set step (3-D coefficient grid_sample
    over the image stack, 144-channel basis per level, decoder 144-64-3, MSE,
    autograd, Adam over all 152 M parameters) on the host cores, on a bounded
    batch of the same workload."""
    import torch
    from ffields import workloads as W
    from oracle import OracleField

    cores = len(os.sched_getaffinity(0))
    torch.set_num_threads(cores)
    N, S = args.dict_images, 256
    cfg = W.dict_cfg(N, S)
    of = OracleField(cfg).requires_grad_(True)
    opt = _oracle_adam(of, cfg)
    n = args.cpu_dict_batch
    g = torch.Generator().manual_seed(W.BASE_SEED + 4001)
    # targets: the generator's images at the drawn pixels
    imgs = torch.stack([W.synth_image(S, seed=W.BASE_SEED + 4000 + i) for i in range(8)])

    def one():
        img = torch.randint(0, 8, (n,), generator=g)
        p = torch.randint(0, S * S, (n,), generator=g)
        x = torch.stack([(p % S).float() + 0.5, (p // S).float() + 0.5, img.float() + 0.5], -1)
        t = imgs[img, p // S, p % S]
        loss = torch.mean((of.field(x) - t) ** 2)
        opt.zero_grad()
        loss.backward()
        opt.step()

    med, k = _time_cpu(one, min_steps=3, max_steps=10)
    return {"value": n / med, "unit": "field pts/s", "cores": cores, "kind": "port",
            "sample": f"{k} timed steps (median) of {n} pixels of the {N}-image workload (images 0-7 as "
                      f"targets), oracle (torch CPU, grid_sample + autograd + Adam over every parameter), "
                      f"{cores} threads on {_cpu_model()}"}


def run_others(args, rank):
    """The SDF, image and dictionary workloads at bounded sizes, each with its
    roofline and CPU baseline, for nesting in the default line."""
    import copy
    import torch
    out = {}
    args = copy.copy(args)
    args.dict_fp32 = args.others_dict_fp32
    legs = [("sdf", lambda: run_sdf(args)[0], cpu_baseline_sdf),
            ("image", lambda: run_image(args), cpu_baseline_image),
            ("dict", lambda: run_dict(args), cpu_baseline_dict)]
    for name, run, cpu in legs:
        log(f"other workload: {name}")
        try:
            r = run()
            keep = {k: r[k] for k in ("value", "unit", "ms_per_step", "steps", "warmup", "dtype", "roofline",
                                      "forward_gather", "config") if k in r}
            if rank == 0 and not args.no_cpu_baseline:
                log(f"cpu baseline: {name}")
                keep["cpu_baseline"] = cpu(args)
            out[name] = keep
        except Exception as e:          # recorded in the line; the headline stays valid
            out[name] = {"error": f"{type(e).__name__}: {e}"}
            log(f"{name} failed: {e}")
        torch.cuda.empty_cache()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="nerf", choices=["sdf", "nerf", "image", "dict", "atomics", "mesh"])
    ap.add_argument("--dict-images", type=int, default=1024)
    ap.add_argument("--dict-batch", type=int, default=1 << 22)
    ap.add_argument("--dict-steps", type=int, default=100, help="training steps before the timed ones")
    ap.add_argument("--dict-fp32", type=int, default=1, help="also train the FP32 form")
    ap.add_argument("--dict-eval-every", type=int, default=10,
                    help="score both forms every this many steps over the second half of training")
    ap.add_argument("--dict-seeds", type=int, default=1,
                    help="2: a second FP32 run on other batches (the FP32 form's run-to-run spread)")
    ap.add_argument("--dict-order", default="cell", choices=["cell", "random"],
                    help="batch order of the dictionary image set (same distribution)")
    ap.add_argument("--nerf-form", default="both", choices=["both", "plain", "proposal"])
    ap.add_argument("--sdf-order", default="morton", choices=["morton", "random"],
                    help="SDF batches: sorted draws from a Morton-ordered pool, or random draw order")
    ap.add_argument("--nerf-pool", type=int, default=1 << 20, help="ground-truth rays the batches are drawn from")
    ap.add_argument("--image-iters", type=int, default=2000)
    ap.add_argument("--nerf-pretrain", type=int, default=1000)
    ap.add_argument("--nerf-test-views", type=int, default=0,
                    help="also score whole held-out views (cameras outside the 100 training views)")
    ap.add_argument("--nerf-test-res", type=int, default=800)
    ap.add_argument("--cpu-batch", type=int, default=1 << 15)
    ap.add_argument("--cpu-rays", type=int, default=256)
    ap.add_argument("--cpu-dict-batch", type=int, default=1 << 14)
    ap.add_argument("--others-dict-fp32", type=int, default=0,
                    help="nested dictionary leg: also train the FP32 form")
    ap.add_argument("--others", type=int, default=1,
                    help="nerf: also run the SDF, image and dictionary workloads (nested in the line)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run three 1 GiB device copies first (reference dispatch for rocprofv3 byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: every GPU box here has one GPU and this build has no "
              "multi-GPU path; run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    rank = int(os.environ.get("RANK", "0"))
    if args.copy_calibration:
        copy_calibration()
    if args.workload == "nerf":
        result, plain = run_nerf(args)
        if rank == 0 and not args.no_cpu_baseline and plain is not None:
            log("cpu baseline")
            result["cpu_baseline"] = cpu_baseline_nerf(args, plain)
        del plain
        if args.others:
            result["config"]["other_workloads"] = run_others(args, rank)
        print(json.dumps(result), flush=True)
        return
    if args.workload == "mesh":
        print(json.dumps(run_mesh(args)), flush=True)
        return
    if args.workload == "atomics":
        result, _ = run_atomics(args)
        print(json.dumps(result), flush=True)
        return
    if args.workload == "image":
        result = run_image(args)
        if rank == 0 and not args.no_cpu_baseline:
            result["cpu_baseline"] = cpu_baseline_image(args)
        print(json.dumps(result), flush=True)
        return
    if args.workload == "dict":
        result = run_dict(args)
        if rank == 0 and not args.no_cpu_baseline:
            result["cpu_baseline"] = cpu_baseline_dict(args)
        print(json.dumps(result), flush=True)
        return
    result, _ = run_sdf(args)
    if rank == 0 and not args.no_cpu_baseline:
        log("cpu baseline")
        result["cpu_baseline"] = cpu_baseline_sdf(args)
    if rank == 0:
        print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
