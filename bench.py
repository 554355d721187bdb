"""rcontrib Mrays/s on one MI355X: shoebox room + 200 tessellated boxes
(25,200 quads = 50,400 triangles), 2,500 sensors, -I+ -ab 4 -ad 4096
-lw 1e-4 -aa 0 -c 1, Reinhart MF:1 sky (145 bins) + ground (1 bin).

A step is one full rcontrib pass over the 2,500 sensors with the scene, its
BVH, the sensor rays and the output records resident on the GPU
(rg_contrib_device); the host-buffer rate (rg_contrib, PCIe copies inside) is
reported beside it.  Rays are ray segments (closest-hit queries),
counted on the device.  Also reported: the DC x sky product
(2500x146 by 146x8760, RGB) on MFMA, the dominant kernel's roofline figures,
and the reference rcontrib timed on the host cores over a bounded sample.
The line also carries gendaymtx (annual sky matrix), rpict (a picture of the
same room) and rfluxmtx (a view matrix of the same room) measurements, each
with its reference program timed beside it.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu-baseline]
                    [--no-batched] [--bvh auto|sah|lbvh]
    python bench.py --only sky|rpict|rflux|config5 [--no-cpu-baseline]

roofline.traffic (HBM bytes per k_trace launch) is read from the committed
rocprofv3 FETCH_SIZE / WRITE_SIZE summary of this workload,
profiles/k_trace_config2_hbm_traffic.md, when it names the same BVH builder.
"""

from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

from pyradiance_amd import _lib, scenes, tune  # noqa: E402

METRIC = "rcontrib Mrays/s on 1 MI355X, plain vs batched, at 3\u03c3 match; s per grad step"
HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
BF16_PEAK_TFLOPS = 2500.0      # MI355X dense BF16 MFMA

# Algorithmic bytes per ray segment: ray record in + hit out (48 B), one node
# per BVH node visited (64 B binary, 128 B four-wide, as the renderer
# reports), one 48-byte primitive record per primitive tested, one 32-byte
# path-state read-modify-write.
B_RAY, B_NODE, B_PRIM, B_STATE = 48.0, 64.0, 48.0, 32.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cores():
    """CPUs this process may run on: the affinity mask, capped by a cgroup
    CPU quota when one is set.  Returns (usable, in the mask, quota or None)."""
    n = len(os.sched_getaffinity(0))
    quota = None
    try:
        q, per = Path("/sys/fs/cgroup/cpu.max").read_text().split()[:2]
        if q != "max":
            quota = max(1, -(-int(q) // int(per)))
    except (OSError, ValueError):
        try:
            q = int(Path("/sys/fs/cgroup/cpu/cpu.cfs_quota_us").read_text())
            per = int(Path("/sys/fs/cgroup/cpu/cpu.cfs_period_us").read_text())
            if q > 0:
                quota = max(1, -(-q // per))
        except (OSError, ValueError):
            pass
    return (min(n, quota) if quota else n), n, quota


def workload(seed: int = 0):
    rad, sensors = scenes.config2(seed)
    mods = []
    m = _lib.RgModifier()
    m.name, m.bin_kind, m.nbins, m.mf, m.rhs = b"skyglow", 1, 145, 1.0, 1.0
    m.nrm[:] = (0.0, 0.0, -1.0)
    m.up[:] = (0.0, 1.0, 0.0)
    g = _lib.RgModifier()
    g.name, g.bin_kind, g.nbins, g.cbin = b"groundglow", 0, 1, 0
    mods = [m, g]
    return rad, sensors, mods


def gpu_params():
    p = _lib.default_params(True)      # rcontrib defaults after override_options
    p.ambounce, p.ambdiv, p.minweight = 4, 4096, 1e-4
    p.seed = 1
    return p


def cpu_baseline(rad: str, sensors: np.ndarray, nsample: int, cores: int, gpu_count=None):
    """Reference rcontrib (oracle/_ref) on `nsample` sensors of the same
    workload, over `cores` processes; gpu_count(sample) -> the GPU's ray
    count on the same sample, reported beside the reference's."""
    ref = ROOT / "oracle" / "_ref"
    if not (ref / "rcontrib").exists() or not (ref / "oconv").exists():
        return None
    cal = ROOT / "oracle" / "reinhart_sky.cal"
    with tempfile.TemporaryDirectory() as td:
        radf = Path(td) / "room.rad"
        octf = Path(td) / "room.oct"
        radf.write_text(rad)
        with open(octf, "wb") as f:
            subprocess.run([str(ref / "oconv"), str(radf)], stdout=f, check=True)
        sample = sensors[:nsample]
        inp = "\n".join(" ".join(f"{v:.17g}" for v in r) for r in sample).encode()
        args = ["-h", "-I+", "-ab", "4", "-ad", "4096", "-lw", "1e-4", "-aa", "0", "-c", "1",
                "-f", str(cal), "-p", "MF=1", "-bn", "Nskybins", "-b", "skybin", "-m", "skyglow",
                "-b", "0", "-bn", "1", "-m", "groundglow"]
        env = dict(os.environ, RAYPATH=f".:{ref / 'lib'}:{ROOT / 'oracle'}")
        # exact ray count of the whole sample: `cores` single-process counting
        # runs over disjoint slices, in parallel (each prints the reference's
        # nrays counter, rt/raytrace.c:606)
        t0 = time.time()
        chunks = np.array_split(np.arange(nsample), cores)
        procs = []
        for k, ch in enumerate(chunks):
            if len(ch) == 0:
                continue
            pf = Path(td) / f"part{k}.txt"
            pf.write_text("\n".join(" ".join(f"{v:.17g}" for v in sample[i]) for i in ch) + "\n")
            fin = open(pf, "rb")
            pr = subprocess.Popen([str(ref / "rcontrib_count"), "-n", "1", *args, str(octf)], stdin=fin,
                                  stdout=subprocess.DEVNULL, stderr=subprocess.PIPE, env=env)
            procs.append((pr, fin))
        nrays = 0
        for pr, fin in procs:
            _, err = pr.communicate()
            fin.close()
            got = [int(ln.split()[1]) for ln in err.decode().splitlines() if ln.startswith("NRAYS")]
            if pr.returncode != 0 or not got:
                return None
            nrays += got[0]
        t_count = time.time() - t0

        def wall(inp_bytes):
            t0 = time.time()
            r = subprocess.run([str(ref / "rcontrib"), "-n", str(cores), *args, str(octf)], input=inp_bytes,
                               capture_output=True, env=env)
            return time.time() - t0, r.returncode
        # all cores over the whole sample, and the octree load alone (a
        # one-sensor run): medians of three
        walls, loads = [], []
        for _ in range(3):
            w, rc = wall(inp)
            if rc != 0:
                return None
            walls.append(w)
            loads.append(wall(inp.split(b"\n")[0] + b"\n")[0])
        w, load = float(np.median(walls)), float(np.median(loads))
        usable, visible, quota = host_cores()
        gr = int(gpu_count(sample)) if gpu_count else None
        return {
            "value": nrays / max(w - load, 1e-9) / 1e6,
            "unit": "Mrays/s",
            "cores": cores,
            "kind": "reference",
            "sample": f"{nsample} of the 2500 sensors, {nsample / cores:.0f} records per process ({nrays} rays, "
                      f"counted exactly by {len(procs)} parallel -n 1 counting runs in {t_count:.1f}s); rcontrib "
                      f"-n {cores} median wall {w:.2f}s of 3 minus median octree load {load:.2f}s of 3 "
                      f"(one-sensor run)",
            "host_cpus": {"affinity": visible, "cgroup_quota": quota, "used": cores},
            "rays_reference": nrays,
            "rays_gpu_same_sample": gr,
            "rays_gpu_over_reference": gr / nrays if gr else None,
        }


def dcsky_full_error(dA, dB, dC, M: int, K: int, N: int, rows: int = 2048):
    """Relative error of every entry of the product dC = dA x dB (per
    channel) against an FP64 product computed on the GPU in row blocks
    (torch float64 matmul), over the entries above 1e-6 of the largest.
    Returns (max, mean, entries)."""
    import torch
    ref_max = 0.0
    for c in range(3):
        Bc = dB[:, :, c].double()
        for r0 in range(0, M, rows):
            ref_max = max(ref_max, (dA[r0:r0 + rows, :, c].double() @ Bc).abs().max().item())
    mx, tot, cnt = 0.0, 0.0, 0
    for c in range(3):
        Bc = dB[:, :, c].double()
        for r0 in range(0, M, rows):
            ref = dA[r0:r0 + rows, :, c].double() @ Bc
            got = dC[r0:r0 + rows, :, c].double()
            nz = ref.abs() > 1e-6 * ref_max
            rel = ((got - ref).abs() / ref.abs().clamp_min(1e-300))[nz]
            if rel.numel():
                mx = max(mx, rel.max().item())
                tot += rel.sum().item()
                cnt += rel.numel()
    return mx, tot / max(cnt, 1), cnt


def dcsky_device_bench(M: int, K: int, N: int, seed: int = 0, reps: int = 5, mode: int | None = None):
    """DC x sky product on device buffers: kernel time (HIP events on the
    null stream, where the product runs), algorithmic and issued MFMA rates,
    and the relative error of every entry against an FP64 product.  A is a
    synthetic coefficient matrix, B the synthetic sky.  mode None: the tile
    form the autotuner picks for these sizes (tune.autotune_dcsky)."""
    import torch
    rng = np.random.default_rng(seed)
    A = (rng.random((M, K, 3), dtype=np.float32) * np.float32(2e-3)).astype(np.float32)
    B = scenes.sky_matrix(seed, K, nhours=N)
    dA = torch.from_numpy(A).cuda()
    dB = torch.from_numpy(B).cuda()
    del A
    dC = torch.empty((M, N, 3), dtype=torch.float32, device="cuda")
    dtuned = None
    if mode is None:
        dtuned = tune.autotune_dcsky(dA.data_ptr(), dB.data_ptr(), dC.data_ptr(), M, K, N)
        mode = dtuned["best"]
    _lib.dcsky_device(dA.data_ptr(), dB.data_ptr(), dC.data_ptr(), M, K, N, mode)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        _lib.dcsky_device(dA.data_ptr(), dB.data_ptr(), dC.data_ptr(), M, K, N, mode)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    emax, emean, nent = dcsky_full_error(dA, dB, dC, M, K, N)
    flops = 2.0 * M * K * N * 3
    byts = 12.0 * (M * K + K * N + M * N)
    del dA, dB, dC
    torch.cuda.empty_cache()
    return {"M": M, "K": K, "N": N, "mode": mode, "kernel_ms": ms, "algorithmic_TFLOPs": flops / ms / 1e9,
            "issued_TFLOPs": 3 * flops / ms / 1e9, "mfma_util_issued": 3 * flops / ms / 1e9 / BF16_PEAK_TFLOPS,
            "GBs": byts / ms / 1e6, "hbm_frac": byts / ms / 1e6 / HBM_PEAK_GBS,
            "max_rel_err_vs_fp64": emax, "mean_rel_err_vs_fp64": emean, "entries_checked": nent,
            "tile_tuning": dtuned}


def committed_traffic(bvh: str, walk: str):
    """HBM bytes per traced ray of the traversal kernel from the committed
    PMC summary, if it was measured with the same BVH builder and walk (its
    machine-readable lines are `bvh: <name>`, `walk: wide|binary` and
    `traffic_bytes_per_ray: <bytes>`).  Per ray, so that it applies to any
    generation window the launch shape uses."""
    f = ROOT / "profiles" / "k_trace_config2_hbm_traffic.md"
    if not f.exists():
        return None
    kv = {}
    for line in f.read_text().splitlines():
        if ":" in line and not line.startswith(("|", " ", "#")):
            k, _, v = line.partition(":")
            kv[k.strip()] = v.strip()
    if kv.get("bvh") != bvh or kv.get("walk", "binary") != walk or "traffic_bytes_per_ray" not in kv:
        return None
    return float(kv["traffic_bytes_per_ray"])


def committed_lane_util():
    """VALU lane utilisation of the persistent traversal kernel from the
    committed SQ counter summary (its machine-readable line
    `lane_utilisation: <fraction>`), or None."""
    f = ROOT / "profiles" / "k_trace_config2_sq_occupancy.md"
    if not f.exists():
        return None
    for line in f.read_text().splitlines():
        if line.startswith("lane_utilisation:"):
            return float(line.partition(":")[2])
    return None


def dcsky_bench(dc: np.ndarray, seed: int = 0):
    nb = dc.shape[1] // 3
    A = dc.reshape(dc.shape[0], nb, 3).astype(np.float32)
    B = scenes.sky_matrix(seed, nb)
    _lib.dcsky(A, B, 1)     # warm-up
    t0 = time.time()
    reps = 5
    for _ in range(reps):
        Cm = _lib.dcsky(A, B, 1)
    wall = (time.time() - t0) / reps
    ref = np.einsum("mkc,knc->mnc", A.astype(np.float64), B.astype(np.float64), optimize=True)
    denom = np.maximum(np.abs(ref), 1e-30)
    nz = np.abs(ref) > 1e-6 * np.abs(ref).max()
    rel = np.abs(Cm - ref)[nz] / denom[nz]
    M, K, N = A.shape[0], nb, B.shape[1]
    return {"M": M, "K": K, "N": N, "ms_host_to_host": wall * 1e3,
            "max_rel_err_vs_fp64": float(rel.max()), "mean_rel_err_vs_fp64": float(rel.mean()),
            "flops": 2.0 * M * K * N * 3}


def copy_calibration(nbytes: int) -> None:
    """One device-to-device copy of `nbytes` (the runtime's plain copy kernel),
    used to scale the FETCH_SIZE / WRITE_SIZE counters to bytes."""
    import ctypes as C
    hip = _lib.hip_runtime()
    a, b = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) or hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)):
        raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, C.c_size_t(nbytes))
    hip.hipDeviceSynchronize()
    hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3)   # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"copy calibration: {nbytes} bytes device to device")


def batched_bench(nsun: int = 255, nscene2: int = 4):
    """Plain (one scene per call) vs batched (one wavefront over many scenes)."""
    res = {}
    # config 1 over many sun positions
    ctxs, rays = [], []
    for seed in range(nsun):
        rad, sens, _ = scenes.config1(seed)
        c = _lib.Context(0)
        c.load_text(rad).commit()
        ctxs.append(c)
        rays.append(sens)
    p = _lib.default_params(False)
    p.ambounce, p.dstrsrc, p.srcsizerat = 0, 0.0, 0.0
    for c in ctxs:
        c.set_params(p)
    _lib.trace_batch(ctxs, rays, _lib.RT_IMM_IRRAD)     # warm-up
    t0 = time.perf_counter()
    nr = 0
    for c, r in zip(ctxs, rays):
        c.trace(r, _lib.RT_IMM_IRRAD)
        nr += c.last_nrays()
    t_plain = time.perf_counter() - t0
    t0 = time.perf_counter()
    _lib.trace_batch(ctxs, rays, _lib.RT_IMM_IRRAD)
    t_batch = time.perf_counter() - t0
    nb = ctxs[0].last_nrays()
    res["config1_suns"] = {"scenes": nsun, "sensors_per_scene": 77, "rays": nb,
                           "plain_ms": t_plain * 1e3, "batched_ms": t_batch * 1e3,
                           "plain_Mrays_s": nr / t_plain / 1e6, "batched_Mrays_s": nb / t_batch / 1e6,
                           "speedup": t_plain / t_batch}
    for c in ctxs:
        c.close()
    # config 2 with several furniture seeds
    ctxs, rays = [], []
    for seed in range(nscene2):
        rad, sens, mods = workload(seed)
        c = _lib.Context(0)
        c.load_text(rad).commit()
        c.set_params(gpu_params())
        ctxs.append(c)
        rays.append(sens)
    t0 = time.perf_counter()
    nr = 0
    for c, r in zip(ctxs, rays):
        c.contrib(r, mods, 1, _lib.RT_IMM_IRRAD)
        nr += c.last_nrays()
    t_plain = time.perf_counter() - t0
    t0 = time.perf_counter()
    _lib.contrib_batch(ctxs, rays, mods, 1, _lib.RT_IMM_IRRAD)
    t_batch = time.perf_counter() - t0
    nb = ctxs[0].last_nrays()
    res["config2_scenes"] = {"scenes": nscene2, "rays": nb, "plain_ms": t_plain * 1e3, "batched_ms": t_batch * 1e3,
                             "plain_Mrays_s": nr / t_plain / 1e6, "batched_Mrays_s": nb / t_batch / 1e6,
                             "speedup": t_plain / t_batch}
    for c in ctxs:
        c.close()
    # config 2's 2,500 sensors as 25 grids of 100 (one call per grid, as a
    # script looping over sensor files would) vs one batched call
    rad, sens, mods = workload(0)
    c = _lib.Context(0)
    c.load_text(rad).commit()
    c.set_params(gpu_params())
    grids = [sens[100 * k:100 * (k + 1)] for k in range(25)]
    c.contrib(grids[0], mods, 1, _lib.RT_IMM_IRRAD)
    t0 = time.perf_counter()
    nr = 0
    for g in grids:
        c.contrib(g, mods, 1, _lib.RT_IMM_IRRAD)
        nr += c.last_nrays()
    t_plain = time.perf_counter() - t0
    t0 = time.perf_counter()
    _lib.contrib_batch([c] * len(grids), grids, mods, 1, _lib.RT_IMM_IRRAD)
    t_batch = time.perf_counter() - t0
    nb = c.last_nrays()
    c.close()
    res["config2_grids"] = {"grids": 25, "sensors_per_grid": 100, "rays": nb, "plain_ms": t_plain * 1e3,
                            "batched_ms": t_batch * 1e3, "plain_Mrays_s": nr / t_plain / 1e6,
                            "batched_Mrays_s": nb / t_batch / 1e6, "speedup": t_plain / t_batch}
    res["config3"] = config3_bench(prims=(1_000, 500_000))
    return res


def config1_bench(nsun: int = 1000, cores: int = 1, cpu: bool = True):
    """Config 1 over `nsun` sun positions (scenes.config1(seed), seeds 0 ..
    nsun-1): rtrace -I -ab 0 -dj 0 -ds 0, 77 sensors per sun.

    GPU, per call: load + commit + trace of one sun's scene, one call after
    another.  GPU, steady state: the committed scenes traced in one batched
    call (rg_trace_batch).  Reference: oconv + rtrace for every sun,
    the suns spread over `cores` concurrent pipelines; per call is the
    median of one sun's oconv + rtrace wall, whole job the wall of all of
    them, trace-only the whole-job wall minus that of the same pipelines
    tracing one sensor each (octree build and load).  Rays: the
    reference's own counter (rtrace_count's NRAYS, rt/raytrace.c:606)
    beside the GPU's, and every irradiance against the reference's."""
    import concurrent.futures as cf
    p = _lib.default_params(False)
    p.ambounce, p.dstrsrc, p.srcsizerat = 0, 0.0, 0.0
    work = [scenes.config1(s) for s in range(nsun)]
    ctxs = [_lib.Context(0) for _ in range(nsun)]
    # per call, sequential
    vals, t_call, g_rays = [], [], []
    for c, (rad, sens, _) in zip(ctxs, work):
        t0 = time.perf_counter()
        c.load_text(rad).commit()
        c.set_params(p)
        v = c.trace(sens, _lib.RT_IMM_IRRAD)
        t_call.append(time.perf_counter() - t0)
        vals.append(v)
        g_rays.append(c.last_nrays())
    # steady state: one batched trace of all the committed scenes
    sens_all = [w[1] for w in work]
    _lib.trace_batch(ctxs, sens_all, _lib.RT_IMM_IRRAD)       # warm-up
    t0 = time.perf_counter()
    out = _lib.trace_batch(ctxs, sens_all, _lib.RT_IMM_IRRAD)
    t_batch = time.perf_counter() - t0
    nb = ctxs[0].last_nrays()
    for i, o in enumerate(out):
        assert np.array_equal(o, vals[i])
    for c in ctxs:
        c.close()
    # rays_gpu: every call fills its scene's shadow cache (as every
    # reference run does); the batched traces reuse the filled caches
    res = {"suns": nsun, "sensors_per_sun": 77, "rays_gpu": int(sum(g_rays)), "rays_gpu_batched": int(nb),
           "gpu_per_call_ms": 1e3 * float(np.mean(t_call)), "gpu_per_call_median_ms": 1e3 * float(np.median(t_call)),
           "gpu_whole_job_s": float(np.sum(t_call)), "gpu_batched_trace_s": t_batch}
    ref = ROOT / "oracle" / "_ref"
    if not cpu or not (ref / "rtrace").exists():
        return res
    env = dict(os.environ, RAYPATH=f".:{ref / 'lib'}")
    args = ["-h", "-I", "-ab", "0", "-dj", "0", "-ds", "0"]
    with tempfile.TemporaryDirectory() as td:
        def job(k, prog="rtrace", nsens=77):
            rad, sens, _ = work[k]
            radf, octf = Path(td) / f"s{k}.rad", Path(td) / f"s{k}_{prog}_{nsens}.oct"
            radf.write_text(rad)
            inp = "\n".join(" ".join(f"{v:.17g}" for v in r) for r in sens[:nsens]).encode() + b"\n"
            t0 = time.perf_counter()
            with open(octf, "wb") as f:
                subprocess.run([str(ref / "oconv"), str(radf)], stdout=f, check=True)
            r = subprocess.run([str(ref / prog), *args, str(octf)], input=inp, capture_output=True, env=env,
                               check=True)
            t = time.perf_counter() - t0
            octf.unlink()
            return t, r.stdout, r.stderr

        def batch(prog, nsens):
            t0 = time.perf_counter()
            with cf.ThreadPoolExecutor(cores) as ex:
                outs = list(ex.map(lambda k: job(k, prog, nsens), range(nsun)))
            return time.perf_counter() - t0, outs
        t_all, outs = batch("rtrace", 77)
        t_load, _ = batch("rtrace", 1)
        _, couts = batch("rtrace_count", 77)
    ref_vals = np.concatenate([np.array(o[1].split(), dtype=np.float64).reshape(-1, 3) for o in outs])
    gpu_vals = np.concatenate(vals)
    rel = np.abs(gpu_vals - ref_vals) / np.maximum(np.abs(ref_vals), 1e-30)
    ref_rays = [int(ln.split()[1]) for o in couts for ln in o[2].decode().splitlines() if ln.startswith("NRAYS")]
    usable, visible, quota = host_cores()
    cpu_call = float(np.median([o[0] for o in outs]))
    # per sun: one call (both sides build the scene and its shadow cache
    # and trace); steady state: the traces alone (GPU: batched over
    # committed scenes; reference: whole-job wall beyond the same pipelines
    # run with one sensor each)
    res.update({
        "cpu_cores": cores, "host_cpus": {"affinity": visible, "cgroup_quota": quota, "used": cores},
        "cpu_per_call_median_ms": 1e3 * cpu_call,
        "cpu_whole_job_s": t_all, "cpu_octree_build_and_load_s": t_load,
        "cpu_whole_job_Mrays_s": sum(ref_rays) / t_all / 1e6,
        "gpu_whole_job_Mrays_s": sum(g_rays) / float(np.sum(t_call)) / 1e6,
        "gpu_steady_ms_per_sun": 1e3 * t_batch / nsun,
        "cpu_steady_ms_per_sun": 1e3 * max(t_all - t_load, 0.0) / nsun,
        "rays_reference": int(sum(ref_rays)),
        "rays_equal": bool(len(ref_rays) == nsun and sum(ref_rays) == sum(g_rays)),
        "rays_equal_per_sun": int(sum(int(a == b) for a, b in zip(ref_rays, g_rays))),
        "max_rel_diff": float(rel.max()), "values": int(rel.size),
        "speedup_per_call": cpu_call / float(np.median(t_call)),
        "speedup_whole_job": t_all / float(np.sum(t_call)),
        "speedup_steady_state": max(t_all - t_load, 0.0) / t_batch,
    })
    return res


def config3_bench(prims=(1_000, 100_000)):
    """64 uneven scenes (rooms 3-20 m, log-uniform primitive and sensor counts,
    own sun each), rcontrib -ab 3 -ad 1024 -aa 0 with MF:1 + ground bins:
    one call per scene vs one batched call over all scenes.  Primitive counts
    are drawn from `prims` (the scene text for 500k-primitive rooms takes
    longer to generate than the whole GPU run)."""
    t0 = time.time()
    sc = scenes.config3(0, prims=prims)
    ctxs, rays = [], []
    _, _, mods = workload(0)
    p = _lib.default_params(True)
    p.ambounce, p.ambdiv, p.seed = 3, 1024, 1
    nprim = 0
    for rad, sens in sc:
        c = _lib.Context(0)
        c.load_text(rad).commit()
        c.set_params(p)
        nprim += c.info()["primitives"]
        ctxs.append(c)
        rays.append(sens)
    t_setup = time.time() - t0
    for c, r in zip(ctxs[:2], rays[:2]):       # warm-up
        c.contrib(r, mods, 1, _lib.RT_IMM_IRRAD)
    t0 = time.perf_counter()
    nr = 0
    per_scene = []
    for c, r in zip(ctxs, rays):
        t1 = time.perf_counter()
        c.contrib(r, mods, 1, _lib.RT_IMM_IRRAD)
        dt = time.perf_counter() - t1
        nr += c.last_nrays()
        per_scene.append({"sensors": int(len(r)), "primitives": int(c.info()["primitives"]),
                          "rays": int(c.last_nrays()), "Mrays_s": c.last_nrays() / dt / 1e6})
    t_plain = time.perf_counter() - t0
    t0 = time.perf_counter()
    _lib.contrib_batch(ctxs, rays, mods, 1, _lib.RT_IMM_IRRAD)
    t_batch = time.perf_counter() - t0
    nb = ctxs[0].last_nrays()
    for c in ctxs:
        c.close()
    return {"scenes": len(sc), "primitives": nprim, "sensors": int(sum(len(r) for r in rays)),
            "prims_range": list(prims), "setup_s": t_setup, "rays_plain": nr, "rays_batched": nb,
            "plain_ms": t_plain * 1e3, "batched_ms": t_batch * 1e3,
            "plain_Mrays_s": nr / t_plain / 1e6, "batched_Mrays_s": nb / t_batch / 1e6,
            "speedup": t_plain / t_batch, "per_scene": per_scene}


def grad_bench(steps: int = 3):
    """Differentiable workload: config-1 room with 24 plastic groups, 1000
    sensors, -ab 3 -ad 2048, binned sky; MSE against targets rendered at the
    true arguments, from a perturbed start.  Times a forward pass and a
    gradient step (forward + path-replay adjoint through torch.autograd)."""
    import torch
    from pyradiance_amd.adjoint import DiffScene, irradiance
    theta_true = scenes.config4_theta(0)
    rad, sens, groups = scenes.config4(theta_true)
    _, _, mods = workload(0)
    p = _lib.default_params(True)
    p.ambounce, p.ambdiv, p.seed = 3, 2048, 1
    sc = DiffScene(rad, sens, groups, mods, p)
    rng = np.random.default_rng(4)
    sky = torch.tensor(scenes.sky_matrix(0, 146, nhours=24)[:, 12], dtype=torch.float32, device="cuda")   # noon
    with torch.no_grad():
        target = irradiance(torch.tensor(theta_true, device="cuda"), sky, sc)
    theta0 = np.clip(theta_true + rng.normal(0, 0.05, theta_true.shape), 0.0, None)
    th = torch.tensor(theta0, device="cuda", requires_grad=True)

    def step():
        e = irradiance(th, sky, sc)
        loss = ((e - target) ** 2).mean()
        loss.backward()
        return float(loss.detach())

    step()
    th.grad = None
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        with torch.no_grad():
            irradiance(th, sky, sc)
    torch.cuda.synchronize()
    t_fwd = (time.perf_counter() - t0) / steps
    nrays = sc.ctx.last_nrays()
    t0 = time.perf_counter()
    for _ in range(steps):
        th.grad = None
        loss = step()
    torch.cuda.synchronize()
    t_grad = (time.perf_counter() - t0) / steps
    gnorm = float(th.grad.norm())
    sc.close()
    return {"sensors": int(len(sens)), "groups": len(groups), "params": 24 * 5 + 146 * 3, "rays_forward": nrays,
            "forward_s": t_fwd, "grad_step_s": t_grad, "grad_over_forward": t_grad / t_fwd, "loss": loss,
            "grad_norm": gnorm}


def config5_bench(seed: int = 0):
    """Config 5 end to end on one GPU: the 10-storey office (2,495,700 quads
    = 5 M triangles, 20,000 sensors) traced with rcontrib -I+ -ab 5 -ad 8192
    -lw 5e-5 -aa 0 into Reinhart MF:4 + ground bins, then that 20000 x 2306
    RGB DC matrix times a synthetic 2306 x 8760 sky on the split-BF16 MFMA
    product."""
    import torch
    t0 = time.perf_counter()
    rad, sens = scenes.config5(seed)
    t_gen = time.perf_counter() - t0
    ctx = _lib.Context(0)
    t0 = time.perf_counter()
    ctx.load_text(rad).commit()
    t_commit = time.perf_counter() - t0
    del rad
    info = ctx.info()
    p = gpu_params()
    p.ambounce, p.ambdiv, p.minweight = 5, 8192, 5e-5
    ctx.set_params(p)
    m = _lib.RgModifier()
    m.name, m.bin_kind, m.nbins, m.mf, m.rhs = b"skyglow", 1, scenes.reinhart_nbins(4), 4.0, 1.0
    m.nrm[:] = (0.0, 0.0, -1.0)
    m.up[:] = (0.0, 1.0, 0.0)
    g = _lib.RgModifier()
    g.name, g.bin_kind, g.nbins, g.cbin = b"groundglow", 0, 1, 0
    mods = [m, g]
    ctx.contrib(sens[::100], mods, 1, _lib.RT_IMM_IRRAD)
    t0 = time.perf_counter()
    out = ctx.contrib(sens, mods, 1, _lib.RT_IMM_IRRAD)
    t_trace = time.perf_counter() - t0
    rays = ctx.last_nrays()
    st = ctx.last_stats()
    ctx.close()
    nb = scenes.reinhart_nbins(4) + 1
    A = np.ascontiguousarray(out.reshape(len(sens), nb, 3), dtype=np.float32)
    del out
    B = scenes.sky_matrix(seed, nb, nhours=8760)
    dA, dB = torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda()
    dC = torch.empty((len(sens), 8760, 3), dtype=torch.float32, device="cuda")
    _lib.dcsky_device(dA.data_ptr(), dB.data_ptr(), dC.data_ptr(), len(sens), nb, 8760, 1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _lib.dcsky_device(dA.data_ptr(), dB.data_ptr(), dC.data_ptr(), len(sens), nb, 8760, 1)
    torch.cuda.synchronize()
    t_dc = time.perf_counter() - t0
    finite = bool(torch.isfinite(dC).all().item())
    return {"workload": "config 5: rcontrib -I+ -ab 5 -ad 8192 -lw 5e-5 -aa 0, Reinhart MF:4 + ground, "
                        "10-storey office, then DC x sky 20000x2306x8760",
            "primitives": int(info["primitives"]), "sensors": int(len(sens)), "scene_generate_s": t_gen,
            "scene_commit_s": t_commit, "rays": int(rays), "trace_s": t_trace, "Mrays_per_s": rays / t_trace / 1e6,
            "rays_per_sensor": rays / len(sens), "walk": "wide" if st.get("node_bytes") == 128 else "binary",
            "dcsky_s": t_dc, "dcsky_output_finite": finite,
            "dc_nonzero_fraction": float((A > 0).mean())}


REF_ENV_PATH = ROOT / "oracle" / "_ref"


def _ref_env():
    ref = REF_ENV_PATH
    return dict(os.environ, RAYPATH=f".:{ref / 'lib'}:{ROOT / 'oracle'}", PATH=f"{ref}:{os.environ.get('PATH', '')}")


def sky_bench(reps: int = 5, cpu: bool = True):
    """gendaymtx on a synthetic annual hourly weather tape (8,760 steps) at
    MF 1 (146 patches) and MF 4 (2,306 patches): host set-up, the device
    matrix (sky patches evaluated and written into HBM), whole gendaymtx()
    calls with float and ASCII output, and the reference gendaymtx on the
    same tape."""
    import torch
    from pyradiance_amd import gen
    wea = scenes.annual_wea(0)
    ref = REF_ENV_PATH / "gendaymtx"
    res = {"workload": "gendaymtx -m N, annual hourly WEA (8760 steps, synthetic, seed 0)"}
    with tempfile.TemporaryDirectory() as td:
        wf = Path(td) / "annual.wea"
        wf.write_text(wea)
        for mf in (1, 4):
            o, _, _ = gen._parse_args(["-m", str(mf)])
            t0 = time.perf_counter()
            sky = _lib.Sky(wea, o)
            plan_ms = (time.perf_counter() - t0) * 1e3
            n = sky.npatch * sky.ncols * 3
            npatch, ncols = sky.npatch, sky.ncols
            out = torch.empty(n, dtype=torch.float32, device="cuda")
            sky.matrix_device(out.data_ptr())
            torch.cuda.synchronize()
            ts = []
            for _ in range(reps):
                t0 = time.perf_counter()
                sky.matrix_device(out.data_ptr())
                torch.cuda.synchronize()
                ts.append(time.perf_counter() - t0)
            sky.close()
            dev_ms = 1e3 * float(np.median(ts))
            t0 = time.perf_counter()
            mine = gen.gendaymtx(str(wf), mfactor=mf, outform="f", header=False)
            tf = (time.perf_counter() - t0) * 1e3
            e = {"patches": npatch, "steps": ncols, "host_setup_ms": plan_ms, "device_call_ms": dev_ms,
                 "device_GB_written_per_s": n * 4 / (dev_ms * 1e-3) / 1e9, "gendaymtx_float_ms": tf,
                 "output_MB": len(mine) / 1e6}
            if mf == 1:
                t0 = time.perf_counter()
                asc = gen.gendaymtx(str(wf), mfactor=mf)
                e["gendaymtx_ascii_ms"] = (time.perf_counter() - t0) * 1e3
            if cpu and ref.exists():
                t0 = time.perf_counter()
                r = subprocess.run([str(ref), "-h", "-m", str(mf), "-of", str(wf)], capture_output=True)
                e["reference_float_ms"] = (time.perf_counter() - t0) * 1e3
                if r.returncode == 0 and len(r.stdout) == len(mine):
                    a = np.frombuffer(mine, dtype=np.float32)
                    b = np.frombuffer(r.stdout, dtype=np.float32)
                    e["max_abs_diff_vs_reference"] = float(np.max(np.abs(a - b)))
                    e["max_rel_diff_vs_reference"] = float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-3)))
                if mf == 1:
                    t0 = time.perf_counter()
                    r = subprocess.run([str(ref), "-m", str(mf), str(wf)], capture_output=True)
                    e["reference_ascii_ms"] = (time.perf_counter() - t0) * 1e3
                    e["ascii_body_identical"] = r.stdout.split(b"\n\n", 1)[-1] == asc.split(b"\n\n", 1)[-1]
            res[f"mf{mf}"] = e
    return res


RPICT_VIEW = ["-vtv", "-vp", "3", "3.8", "1.5", "-vd", "0", "-1", "0", "-vu", "0", "0", "1", "-vh", "80", "-vv", "60"]
RPICT_OPTS = ["-ab", "1", "-ad", "1024", "-aa", "0"]


def rpict_bench(xres: int = 1024, yres: int = 1024, reps: int = 3, cpu: bool = True, cpu_res: int = 128):
    """rpict of the rcontrib room (config-2 scene: room, window, 200 boxes,
    sky and ground glow) from its north side toward the window, one primary
    ray per pixel (-ps 1) with -ab 1 -ad 1024 -aa 0 and rpict's other
    defaults; the reference rpict on a smaller picture of the same view."""
    from pyradiance_amd import params as _params
    from pyradiance_amd import rt as _rt
    from pyradiance_amd import view as _view
    rad, _ = scenes.config2(0)
    with tempfile.TemporaryDirectory() as td:
        radf = Path(td) / "room.rad"
        radf.write_text(rad)
        ctx = _lib.Context(0)
        ctx.load_text(rad).commit()
        opt = _params.parse(RPICT_OPTS, rcontrib=False)
        p = opt.params
        p.srcsizerat, p.maxdepth, p.minweight, p.dstrsrc, p.seed = 0.25, 7, 1e-4, 0.0, 1
        ctx.set_params(p)
        v, _ = _view.parse_view(RPICT_VIEW)
        v.setup()
        _, xr, yr = _view.normaspect(v.aspect(), 1.0, xres, yres)
        ctx.trace_view(v, xr, yr, 0.67)
        ts, rays = [], []
        for _ in range(reps):
            t0 = time.perf_counter()
            img = ctx.trace_view(v, xr, yr, 0.67)
            ts.append(time.perf_counter() - t0)
            rays.append(ctx.last_nrays())
        sec = float(np.median(ts))
        e = {"workload": f"rpict {' '.join(RPICT_VIEW + RPICT_OPTS)} -ps 1 -x {xres} -y {yres} (config-2 room)",
             "pixels": xr * yr, "picture_ms": sec * 1e3, "Mrays_per_s": float(np.mean(rays)) / sec / 1e6,
             "rays_per_pixel": float(np.mean(rays)) / (xr * yr), "Mpixels_per_s": xr * yr / sec / 1e6,
             "mean_radiance": float(img.mean())}
        # the whole rpict() call: options, view, rays, trace, RGBE picture
        _rt.rpict(RPICT_VIEW, radf, xres=xres, yres=yres, params=RPICT_OPTS + ["-ps", "1"])
        t0 = time.perf_counter()
        pic = _rt.rpict(RPICT_VIEW, radf, xres=xres, yres=yres, params=RPICT_OPTS + ["-ps", "1"])
        e["rpict_call_ms"] = (time.perf_counter() - t0) * 1e3
        e["picture_bytes"] = len(pic)
        ref = REF_ENV_PATH
        if cpu and (ref / "rpict").exists() and (ref / "oconv").exists():
            octf = Path(td) / "room.oct"
            with open(octf, "wb") as f:
                subprocess.run([str(ref / "oconv"), str(radf)], stdout=f, check=True)

            def run(x, y):
                t0 = time.perf_counter()
                r = subprocess.run([str(ref / "rpict"), *RPICT_VIEW, *RPICT_OPTS, "-ps", "1", "-x", str(x), "-y",
                                    str(y), str(octf)], capture_output=True, env=_ref_env())
                return time.perf_counter() - t0, r.returncode
            w, rc = run(cpu_res, cpu_res)
            load, rc1 = run(1, 1)
            if rc == 0 and rc1 == 0:
                _, cx, cy = _view.normaspect(v.aspect(), 1.0, cpu_res, cpu_res)
                e["cpu_baseline"] = {"value": cx * cy / max(w - load, 1e-9) / 1e6, "unit": "Mpixels/s", "cores": 1,
                                     "kind": "reference",
                                     "sample": f"{cx}x{cy} picture of the same view and options, wall {w:.2f}s "
                                               f"minus a 1x1 run {load:.2f}s (octree load)"}
    return e


RFLUX_OPTS = ["-I+", "-ab", "4", "-ad", "4096", "-lw", "1e-4", "-aa", "0"]


def _window_receiver(room) -> str:
    """Klems full receiver just inside the window, facing the room."""
    x0 = 0.5 * (room.width - room.win_w)
    x1, z0 = x0 + room.win_w, room.win_sill
    z1 = z0 + room.win_h
    return ("#@rfluxmtx h=kf u=Z\nvoid glow winglow\n0\n0\n4 1 1 1 0\n\n"
            f"winglow polygon window_recv\n0\n0\n12 {x0} .01 {z0} {x0} .01 {z1} {x1} .01 {z1} {x1} .01 {z0}\n")


def rflux_bench(reps: int = 3, cpu: bool = True, cpu_sample: int = 1000, cores: int = 16):
    """rfluxmtx view matrix: the 2,500 sensors of the rcontrib room (room and
    200 boxes, no sky) to a Klems full (145 bin) receiver on the window,
    -I+ -ab 4 -ad 4096 -lw 1e-4 -aa 0, ASCII output; the whole rfluxmtx()
    call (receiver parsing, rcontrib options, trace, formatting) and the
    reference rfluxmtx (forking the reference rcontrib) on a sample."""
    from pyradiance_amd import rflux
    rng = np.random.default_rng(0)
    room = scenes.Room()
    rad = scenes.room_rad(room) + scenes.furniture_rad(rng, 200, room, 3, 7)
    _, sensors = scenes.config2(0)
    text = "".join(" ".join(f"{v:.17g}" for v in r) + "\n" for r in sensors).encode()
    with tempfile.TemporaryDirectory() as td:
        radf, recf = Path(td) / "room.rad", Path(td) / "window.rad"
        radf.write_text(rad)
        recf.write_text(_window_receiver(room))
        rflux.rfluxmtx(recf, rays=text, params=RFLUX_OPTS, scene=[radf])
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            out = rflux.rfluxmtx(recf, rays=text, params=RFLUX_OPTS, scene=[radf])
            ts.append(time.perf_counter() - t0)
        sec = float(np.median(ts))
        body = out.split(b"\n\n", 1)[-1]
        vals = np.array(body.split(), dtype=np.float64)
        e = {"workload": f"rfluxmtx {' '.join(RFLUX_OPTS)} - window.rad(h=kf) room.rad, 2500 sensors",
             "records": len(sensors), "call_ms": sec * 1e3, "records_per_s": len(sensors) / sec,
             "values": int(vals.size), "mean_value": float(vals.mean()) if vals.size else None}
        ref = REF_ENV_PATH
        if cpu and (ref / "rfluxmtx").exists() and (ref / "rcontrib").exists():
            def run(k):
                t0 = time.perf_counter()
                r = subprocess.run([str(ref / "rfluxmtx"), "-n", str(cores), *RFLUX_OPTS, "-", str(recf), str(radf)],
                                   input=b"".join(text.splitlines(keepends=True)[:k]), capture_output=True,
                                   env=_ref_env(), cwd=td)
                return time.perf_counter() - t0, r.returncode
            w, rc = run(cpu_sample)
            load, rc1 = run(1)
            if rc == 0 and rc1 == 0:
                e["cpu_baseline"] = {"value": cpu_sample / max(w - load, 1e-9), "unit": "records/s",
                                     "cores": cores, "kind": "reference",
                                     "sample": f"{cpu_sample} of the 2500 sensors, rfluxmtx -n {cores} wall {w:.2f}s "
                                               f"minus a one-sensor run {load:.2f}s (oconv and octree load)"}
    return e


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-grad", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shape")
    ap.add_argument("--shape", default="", metavar="K=V,...",
                    help="launch shape to use instead of autotuning, e.g. trace_kernel=1,refill=16")
    ap.add_argument("--no-config5", action="store_true",
                    help="skip config 5 (the 20000x2306x8760 DC x sky product and the 5 M-triangle rcontrib pass)")
    ap.add_argument("--no-sky", action="store_true", help="skip the gendaymtx measurement")
    ap.add_argument("--no-rpict", action="store_true", help="skip the rpict measurement")
    ap.add_argument("--no-rflux", action="store_true", help="skip the rfluxmtx measurement")
    ap.add_argument("--only", choices=["sky", "rpict", "rflux", "config1", "config5"], default=None,
                    help="run one of the gendaymtx / rpict / rfluxmtx measurements alone and print its JSON")
    ap.add_argument("--cpu-sample", type=int, default=0,
                    help="config-2 sensors timed on the CPU (default: 20 records per host core, at least 480)")
    ap.add_argument("--no-config1", action="store_true", help="skip the 1000-sun config-1 comparison")
    ap.add_argument("--sensors", type=int, default=2500)
    ap.add_argument("--bvh", choices=["auto", "sah", "lbvh"], default="auto")
    ap.add_argument("--max-big", type=int, default=None, help="room-sized primitives kept out of the BVH")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="also run one device copy of BYTES (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the forward pass shards "
              "by sensor rows with no communication; this measurement runs on one GPU)", file=sys.stderr)
        sys.exit(2)

    if args.only:
        if args.copy_calibration:
            copy_calibration(args.copy_calibration)
        cpu_leg = not args.no_cpu_baseline
        fn = {"sky": lambda: sky_bench(cpu=cpu_leg), "rpict": lambda: rpict_bench(cpu=cpu_leg),
              "rflux": lambda: rflux_bench(cpu=cpu_leg, cores=host_cores()[0]),
              "config1": lambda: config1_bench(cores=host_cores()[0], cpu=cpu_leg),
              "config5": config5_bench}[args.only]
        print(json.dumps(fn()))
        return

    rad, sensors, mods = workload(0)
    sensors = sensors[: args.sensors]
    ctx = _lib.Context(0)
    ctx.set_bvh_builder({"auto": _lib.BVH_AUTO, "sah": _lib.BVH_SAH, "lbvh": _lib.BVH_LBVH}[args.bvh])
    if args.max_big is not None:
        ctx.set_tuning(max_big=args.max_big)
    t0 = time.time()
    ctx.load_text(rad).commit()
    t_commit = time.time() - t0
    info = ctx.info()
    log(f"scene: {info}, commit+BVH {t_commit * 1e3:.1f} ms")
    ctx.set_params(gpu_params())
    flags = _lib.RT_IMM_IRRAD
    scene_tuned = None
    if not args.no_autotune and args.bvh == "auto" and args.max_big is None:
        # tree construction (builder, room-sized primitives kept out of the
        # tree) timed on the workload itself, host buffers, before the
        # launch shape
        scene_tuned = tune.autotune_scene(ctx, ctx.commit, lambda: ctx.contrib(sensors, mods, 1, flags))
        log(f"scene autotune: best {scene_tuned['best']} {scene_tuned['ms']:.1f} ms; "
            f"tried {[(c, round(m, 1), round(cm, 1)) for c, m, cm in scene_tuned['tried']]}")
    # sensors and records resident in HBM for the timed steps (rg_contrib_device)
    import torch
    d_sens = torch.from_numpy(np.ascontiguousarray(sensors, dtype=np.float64)).cuda()
    ncol = 3 * sum(m.nbins for m in mods)
    d_out = torch.empty((len(sensors), ncol), dtype=torch.float64, device="cuda")

    def step():
        ctx.contrib_device(d_sens.data_ptr(), len(sensors), mods, d_out.data_ptr(), 1, flags)

    tuned = None
    if args.shape:
        shape = {k: int(v) for k, v in (kv.split("=") for kv in args.shape.split(","))}
        ctx.set_tuning(**shape)
        tuned = {"best": dict(tune.DEFAULT, **shape), "ms": None, "tried": [(None, None)]}
    elif not args.no_autotune:
        tuned = tune.autotune(ctx, step)
        top = sorted(tuned["tried"], key=lambda x: x[1])[:5]
        log(f"autotune: best {tuned['best']} {tuned['ms']:.1f} ms; per-thread kernel "
            f"{tuned['tried'][0][1]:.1f} ms; top {[(t, round(m, 1)) for t, m in top]}")

    for i in range(args.warmup):
        step()
        log(f"warmup {i} done: {ctx.last_nrays()} rays")
    times, rays = [], []
    # HIP events around every traversal and shading launch, on the context's
    # stream, during the timed steps (the renderer synchronises once per
    # window anyway, to size the next one, so the events add no waits)
    ctx.set_profiling(1)
    st_t = {"trace_ms": 0.0, "shade_ms": 0.0, "launches": 0.0, "rays": 0.0}
    torch.cuda.synchronize()
    for i in range(args.steps):
        t0 = time.perf_counter()
        step()
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
        rays.append(ctx.last_nrays())
        st = ctx.last_stats()
        for k in st_t:
            st_t[k] += st[k]
        log(f"step {i}: {times[-1] * 1e3:.1f} ms, {rays[-1]} rays")
    ctx.set_profiling(0)
    nsteps = max(args.steps, 1)
    ms = 1e3 * float(np.mean(times))
    value = float(np.sum(rays)) / float(np.sum(times)) / 1e6
    out = d_out.cpu().numpy()
    # the same step through the host-buffer entry point (ray upload and
    # record download over PCIe inside the call), for reference
    t0 = time.perf_counter()
    ctx.contrib(sensors, mods, 1, flags)
    host_rate = ctx.last_nrays() / (time.perf_counter() - t0) / 1e6

    # node and primitive counts per ray: a separate counting run
    ctx.set_profiling(2)
    step()
    st_c = ctx.last_stats()
    ctx.set_profiling(0)
    nr = st_c["rays"]
    b_node = st_c.get("node_bytes") or B_NODE   # 64 B binary node, 128 B four-wide node
    per_ray = B_RAY + b_node * st_c["nodes"] / nr + B_PRIM * st_c["prims"] / nr + B_STATE
    avg_launch_ms = st_t["trace_ms"] / max(st_t["launches"], 1)
    bytes_per_launch = per_ray * st_t["rays"] / max(st_t["launches"], 1)
    achieved = bytes_per_launch / (avg_launch_ms * 1e-3) / 1e9
    log(f"k_trace: {st_t['launches'] / nsteps:.0f} launches, {st_t['trace_ms'] / nsteps:.1f} ms per step; "
        f"k_shade {st_t['shade_ms'] / nsteps:.1f} ms; "
        f"{st_c['nodes'] / nr:.1f} nodes, {st_c['prims'] / nr:.1f} prims per ray; {per_ray:.0f} B/ray")

    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
    dcs = dcsky_bench(out)
    dcs["device"] = dcsky_device_bench(2500, 146, 8760)
    log(f"dcsky: {dcs}")
    dc5 = c5 = None
    if not args.no_config5:
        dc5 = dcsky_device_bench(20000, 2306, 8760)
        log(f"dcsky config 5 size: {dc5}")
        c5 = config5_bench()
        log(f"config 5 end to end: {c5}")

    batched = None
    if not args.no_batched:
        batched = batched_bench()
        log(f"batched: { {k: v for k, v in batched.items() if k != 'config3'} }")
        c3 = {k: v for k, v in batched["config3"].items() if k != "per_scene"}
        log(f"config3: {c3}")

    grad = None
    if not args.no_grad:
        grad = grad_bench()
        log(f"grad step: {grad}")

    sky = None
    if not args.no_sky:
        sky = sky_bench(cpu=not args.no_cpu_baseline)
        log(f"gendaymtx: {sky}")
    pict = None
    if not args.no_rpict:
        pict = rpict_bench(cpu=not args.no_cpu_baseline)
        log(f"rpict: {pict}")
    rfx = None
    if not args.no_rflux:
        rfx = rflux_bench(cpu=not args.no_cpu_baseline, cores=host_cores()[0])
        log(f"rfluxmtx: {rfx}")
    c1 = None
    if not args.no_config1:
        c1 = config1_bench(cores=host_cores()[0], cpu=not args.no_cpu_baseline)
        log(f"config 1 over 1000 suns: {c1}")

    cpu = None
    if not args.no_cpu_baseline:
        cores = host_cores()[0]
        nsample = min(len(sensors), args.cpu_sample or max(480, 20 * cores))

        def gpu_count(sample):
            ctx.contrib(sample, mods, 1, flags)
            return ctx.last_nrays()
        try:
            cpu = cpu_baseline(rad, sensors, nsample, cores, gpu_count)
        except Exception as e:  # the GPU figure stands on its own
            log(f"cpu baseline failed: {e}")

    bvh_name = args.bvh if args.bvh != "auto" else ("sah" if info["primitives"] <= (4 << 20) else "lbvh")
    if scene_tuned:
        bvh_name = "sah" if scene_tuned["best"]["builder"] == _lib.BVH_SAH else "lbvh"
    walk = "wide" if st_c.get("node_bytes") == 128 else "binary"
    traffic_ray = committed_traffic(bvh_name, walk)
    # HBM bytes per launch = measured bytes per ray x rays per launch of this run
    traffic = traffic_ray * st_t["rays"] / max(st_t["launches"], 1) if traffic_ray else None
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 geometry / f32 BVH+colour",
        "data": "synthetic (seeded scene generator)",
        "config": {"workload": "rcontrib -I+ -ab 4 -ad 4096 -lw 1e-4 -aa 0 -c 1, Reinhart MF:1 + ground",
                   "sensors": int(len(sensors)), "primitives": int(info["primitives"]),
                   "rays_per_step": int(np.mean(rays)), "scene_commit_ms": t_commit * 1e3,
                   "io": "sensors and records resident in HBM (rg_contrib_device)",
                   "host_buffer_Mrays_s": host_rate,
                   "bvh": bvh_name,
                   "launch_shape": tuned["best"] if tuned else tune.DEFAULT,
                   "tree": scene_tuned["best"] if scene_tuned else None,
                   "per_thread_kernel_ms": tuned["tried"][0][1] if tuned else None},
        # bound: the roofline the algorithmic bytes are priced against;
        # limiter: what the counters show limits the kernel (instruction
        # issue with a third of the lanes active: DRAM traffic is a few
        # percent of HBM peak, dram_frac)
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "limiter": "issue",
                     "dram_frac": (traffic / (avg_launch_ms * 1e-3) / 1e9 / HBM_PEAK_GBS) if traffic else None,
                     "lane_utilisation": committed_lane_util(),
                     "kernel": "k_trace", "avg_launch_ms": avg_launch_ms,
                     "shade_kernel_ms_per_step": st_t["shade_ms"] / nsteps,
                     "trace_kernel_ms_per_step": st_t["trace_ms"] / nsteps,
                     "launches_timed": int(st_t["launches"]),
                     "bytes_per_ray": per_ray, "nodes_per_ray": st_c["nodes"] / nr, "walk": walk,
                     "prims_per_ray": st_c["prims"] / nr},
        "dcsky": dcs,
        "dcsky_config5": dc5,
        "config5": c5,
        "batched": batched,
        "grad_step": grad,
        "gendaymtx": sky,
        "rpict": pict,
        "rfluxmtx": rfx,
        "config1": c1,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
