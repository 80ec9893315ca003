"""Benchmark of the marching-cubes isosurface path on one MI355X.

Workload: a 1024^3 float32 turbulence-like field (sum of 64 random-phase
Fourier modes, |k| in [1, 64], amplitude |k|^(-5/3), normalised to [-1, 1]),
isolevels {-0.3, 0.0, 0.3} extracted together, VertexNormal (shared vertices,
per-vertex normals).  One step = one kmc_extract_levels call on the
HBM-resident volume with device-resident outputs.  The volume min/max is
computed once beforehand, as the reference caches it in the volume object.

Printed: one JSON line with Gvoxels/s (voxel classifications per second,
N * levels / t), Mtriangles/s, the HBM roofline of the job and the
reference's own CPU path timed on a z-slab sample of the same volume.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--size 1024] [--no-cpu-baseline]
"""
import argparse
import ctypes
import json
import math
import os
import platform
import sys
import threading
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
LEVELS = (-0.3, 0.0, 0.3)
METRIC = "Gvoxels/sec + Mtriangles/sec, % of HBM roofline; batched vs plain speedup; p50/p99"


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--size", type=int, default=1024)
    ap.add_argument("--cpu-slices", type=int, default=192)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--profile-steps", type=int, default=3, help="extra steps timed kernel by kernel")
    ap.add_argument("--batch-jobs", type=int, default=256, help="volumes of the batched leg (0: skip it)")
    ap.add_argument("--batch-steps", type=int, default=5)
    ap.add_argument("--serve-volumes", type=int, default=32, help="resident volumes of the serving leg (0: skip it)")
    ap.add_argument("--serve-requests", type=int, default=300)
    ap.add_argument("--serve-workers", type=int, default=0, help="server worker threads (0: the library's default)")
    ap.add_argument("--serve-max-batch", type=int, default=0,
                    help="most requests one server batch takes (0: the library's default)")
    ap.add_argument("--no-ct", action="store_true", help="skip the 512^3 uint8 CT-like leg")
    ap.add_argument("--no-rows", action="store_true",
                    help="skip the SlicePlane / unstructured / KVSML / STL leg")
    ap.add_argument("--only-serving", action="store_true",
                    help="only the serving leg (its plain calls, then the server): for profiling")
    ap.add_argument("--only-batched-ct", action="store_true",
                    help="only the batched and the 512^3 CT legs (no CPU reference): for profiling")
    ap.add_argument("--only-main", action="store_true",
                    help="only the 1024^3 step (no CPU baseline, rows, batched, serving, CT or FP64 legs): for profiling")
    args = ap.parse_args()
    if args.only_main:
        args.no_cpu_baseline, args.no_ct, args.no_rows, args.batch_jobs, args.serve_volumes = True, True, True, 0, 0
    if args.gpus != 1:
        print("bench.py measures one GPU only (across GPUs the path would shard by z-slab); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from kvs_amd import datasets, kmc

    lib = kmc.load()
    dev = torch.device("cuda:0")
    if args.only_serving:
        torch.cuda.init()
        print(json.dumps({"serving": serving_leg(lib, kmc, datasets, torch, args)}), flush=True)
        return
    if args.only_batched_ct:
        torch.cuda.init()
        args.no_cpu_baseline = True
        print(json.dumps({"batched": batched_leg(lib, kmc, datasets, torch, args),
                          "ct_512": ct_leg(lib, kmc, datasets, torch, args)}), flush=True)
        return
    n = args.size
    print(f"[bench] generating {n}^3 float32 turbulence field on the GPU", file=sys.stderr, flush=True)
    t0 = time.time()
    vol = datasets.turbulence_torch(n, seed=2, device=dev)
    torch.cuda.synchronize()
    print(f"[bench] generated in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)

    stream = torch.cuda.current_stream(dev)
    sptr = ctypes.c_void_p(stream.cuda_stream)
    copy = copy_ceiling(torch, dev)
    v = kmc.make_volume(vol.data_ptr(), dims=(n, n, n), memory=kmc.MEMORY_DEVICE, dtype=np.float32)
    mn, mx = ctypes.c_double(), ctypes.c_double()
    st = lib.kmc_volume_min_max(ctypes.byref(v), sptr, ctypes.byref(mn), ctypes.byref(mx))
    assert st == 0, st
    v.has_min_max_values = 1
    v.min_value, v.max_value = mn.value, mx.value

    nlev = len(LEVELS)
    levels = (ctypes.c_double * nlev)(*LEVELS)
    outs = (kmc.Surface * nlev)()

    def step():
        s = lib.kmc_extract_levels(ctypes.byref(v), levels, nlev, kmc.VERTEX_NORMAL, 0, kmc.MEMORY_DEVICE,
                                   sptr, outs)
        if s != 0:
            raise kmc.KmcError(s)
        res = [(int(o.n_vertices), int(o.n_triangles)) for o in outs]
        for i in range(nlev):
            lib.kmc_surface_release(ctypes.byref(outs[i]))
        return res

    print("[bench] warmup", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        counts = step()
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    print(f"[bench] timing {args.steps} steps", file=sys.stderr, flush=True)
    wall0 = time.perf_counter()
    start.record(stream)
    for _ in range(args.steps):
        counts = step()
    end.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    ms = start.elapsed_time(end) / args.steps

    nodes = n ** 3
    ntri = sum(t for _, t in counts)
    nvert = sum(vv for vv, _ in counts)
    sec = ms / 1e3
    gvox = nodes * nlev / sec / 1e9
    mtri = ntri / sec / 1e6
    # algorithmic bytes of the job: volume read once + returned arrays
    # (VertexNormal: coords 12 B and normals 12 B per vertex, connections 12 B per triangle)
    b_alg = nodes * 4 + sum(24 * vv + 12 * t for vv, t in counts)
    achieved = b_alg / sec / 1e9

    # per-kernel device time (HIP events between the launches, on the same stream)
    lib.kmc_set_timing(1)
    prof = {"classify": 0.0, "count": 0.0, "scan": 0.0, "compact": 0.0, "vertices": 0.0, "triangles": 0.0,
            "vertex_normals": 0.0, "emit": 0.0}
    for _ in range(args.profile_steps):
        step()
        t = kmc.Timing()
        lib.kmc_last_timing(ctypes.byref(t))
        for k in prof:
            prof[k] += getattr(t, k + "_ms") / args.profile_steps
    lib.kmc_set_timing(0)
    kernels = kernel_rooflines(prof, n, nlev, counts)

    cpu = fp64 = None
    if not args.no_cpu_baseline:
        cpu, fp64 = cpu_baseline(vol, args.cpu_slices, mn.value, mx.value, kmc)
    rows = None
    if not args.no_rows:
        rows = rows_leg(lib, kmc, datasets, torch, vol, v, mn.value, mx.value, sptr, not args.no_cpu_baseline)
    del vol
    torch.cuda.empty_cache()

    batched = None
    if args.batch_jobs > 0:
        batched = batched_leg(lib, kmc, datasets, torch, args)
        torch.cuda.empty_cache()

    serving = None
    if args.serve_volumes > 0:
        serving = serving_leg(lib, kmc, datasets, torch, args)
        torch.cuda.empty_cache()

    ct = None
    if not args.no_ct:
        ct = ct_leg(lib, kmc, datasets, torch, args)

    out = {
        "metric": METRIC,
        "value": round(gvox, 3),
        "unit": "Gvoxels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"{n}^3 float32 turbulence field (64 Fourier modes, |k|^-5/3), isolevels "
                        f"{list(LEVELS)} fused, VertexNormal, device-resident volume and outputs",
            "resolution": [n, n, n], "levels": list(LEVELS), "normal_type": "VertexNormal",
            "duplication": False,
        },
        "mtriangles_per_s": round(mtri, 2),
        "triangles_per_step": ntri,
        "vertices_per_step": nvert,
        "wall_ms_per_step": round(wall / args.steps * 1e3, 3),
        "roofline": dominant_roofline(kernels, n, 4),
        "kernels": kernels,
        "job_roofline": {
            "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "algorithmic_bytes": b_alg,
            "scope": "whole kmc_extract_levels call (classify, count, scan, host sync, emit): volume read "
                     "once + returned arrays, over ms_per_step",
        },
        "cpu_baseline": cpu,
        "copy_ceiling": copy,
        "batched": batched,
        "serving": serving,
        "ct_512": ct,
        "fp64_reevaluation": fp64,
        "rows": rows,
    }
    print(json.dumps(out), flush=True)


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "kmc_turbulence1024_f32_vn3_traffic.md")


def measured_traffic():
    """HBM bytes per launch of each kernel of this workload, as measured by the
    two rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE) recorded in the profile
    summary (its json block), or {} if there is none."""
    try:
        text = open(TRAFFIC_PROFILE).read()
    except OSError:
        return {}
    if "```json" not in text:
        return {}
    block = text.split("```json", 1)[1].split("```", 1)[0]
    return json.loads(block).get("kernels", {})


def kernel_rooflines(prof, n, nlev, counts):
    """Algorithmic bytes and achieved bandwidth of each kernel of one step
    (one launch each; the emit stage is compact + vertices + triangles +
    vertex_normals).  Per launch:
      classify        reads the volume (4 B per node), writes the classification
                      masks (nlev bits per node, in 64-node words);
      count           reads the masks, writes 2 x 4 B per node row and level and
                      the in-row chunk prefixes (4 B per 64-node chunk and level
                      for rows of at most 64 chunks, else 2 x 4 B);
      scan            reads those row counts, writes 8 B prefixes;
      compact         reads the chunk prefixes (as written by count) and the
                      masks, writes the records
                      (8 B per vertex, 8 + 1 B per triangle; the 32-byte crossing
                      records of chunks with a vertex are not counted);
      vertices        reads each vertex key and its edge's two values, writes its
                      coordinates: 8 + 8 + 12 B per vertex;
      triangles       reads each record, cube index and its three vertices'
                      coordinates, writes connections and face normal:
                      8 + 1 + 36 + 12 + 12 B per triangle (the cell-slot writes,
                      about 8 B per triangle, are not counted);
      vertex_normals  reads each vertex key, its four cell slots and the cells'
                      cube indices, and the face normals of the triangles using it
                      (3 per triangle, 12 B each), writes the normal: 8 + 16 + 4 + 12
                      B per vertex + 36 B per triangle."""
    nc = (n + 63) // 64
    masks = nlev * n * n * nc * 8
    chunks = nlev * n * n * nc
    prefix_bytes = (4 if nc <= 64 else 8) * chunks
    rows = n * n
    nv = sum(v for v, _ in counts)
    nt = sum(t for _, t in counts)
    algo = {
        "classify": n ** 3 * 4 + masks,
        "count": masks + 2 * nlev * rows * 4 + prefix_bytes,
        "scan": 2 * nlev * rows * (4 + 8),
        "compact": prefix_bytes + masks + 8 * nv + 9 * nt,
        "vertices": 28 * nv,
        "triangles": 69 * nt,
        "vertex_normals": 40 * nv + 36 * nt,
    }
    names = {"classify": "k_classify<float, 3>", "count": "k_count", "scan": "k_scan_*", "compact": "k_compact",
             "vertices": "k_vertices<float>", "triangles": "k_triangles<float, 2>",
             "vertex_normals": "k_vertex_normals"}
    traffic = measured_traffic()
    out = {}
    for k in names:
        ms = prof[k]
        gbs = algo[k] / (ms / 1e3) / 1e9 if ms > 0 else 0.0
        out[k] = {"kernel": names[k], "ms": round(ms, 4), "algorithmic_bytes": algo[k], "achieved_gbs": round(gbs, 1),
                  "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": traffic.get(names[k])}
    out["emit_total_ms"] = round(prof["emit"], 4)
    return out


def dominant_roofline(kernels, n, itemsize):
    """The longest kernel of the step (classify) against HBM peak.  Its share
    of the job's algorithmic bytes is the volume read once (N * sizeof(T));
    the classification masks it writes are intermediate data, reported apart
    and left out of `achieved`."""
    k = max((x for x in kernels if isinstance(kernels[x], dict)), key=lambda x: kernels[x]["ms"])
    e = kernels[k]
    alg = n ** 3 * itemsize if k == "classify" else e["algorithmic_bytes"]
    gbs = alg / (e["ms"] / 1e3) / 1e9 if e["ms"] > 0 else 0.0
    return {"bound": "hbm", "kernel": e["kernel"], "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": e["traffic"], "algorithmic_bytes": alg,
            "intermediate_bytes": e["algorithmic_bytes"] - alg, "launch_ms": e["ms"],
            "scope": "this kernel alone; the whole call's fraction of HBM peak is job_roofline.frac"}


def copy_ceiling(torch, dev, nbytes=1 << 30, reps=5):
    """Achievable HBM ceiling: a plain device-to-device copy of 1 GiB (torch's
    copy kernel), read + write bytes over its event-timed duration.  Its
    FETCH_SIZE / WRITE_SIZE counters also calibrate the traffic counters."""
    src = torch.empty(nbytes // 4, dtype=torch.float32, device=dev).fill_(1.0)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        dst.copy_(src)
    b.record()
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / reps
    del src, dst
    return {"bytes_per_copy": nbytes, "ms": round(ms, 4), "gbs": round(2 * nbytes / (ms / 1e3) / 1e9, 1)}


def batched_leg(lib, kmc, datasets, torch, args):
    """Batched vs plain on the same GPU, both on HBM-resident volumes with
    device-resident outputs: `plain` is one kmc_extract per (volume, isolevel),
    in sequence (each call returns when its surface is complete); `batched` is
    one run of a planned kmc_batch (one graph launch) over all of them."""
    import numpy as np

    jobs = datasets.batch_workload(args.batch_jobs)
    print(f"[bench] batched leg: generating {len(jobs)} blob volumes on the GPU", file=sys.stderr, flush=True)
    vols, specs = [], []
    for dims, seed, nb, levels in jobs:
        t = datasets.blobs_torch(dims, seed=seed, nblobs=nb, device="cuda")
        v = kmc.make_volume(t.data_ptr(), dims=dims, memory=kmc.MEMORY_DEVICE, dtype=np.float32)
        mn, mx = ctypes.c_double(), ctypes.c_double()
        torch.cuda.synchronize()
        assert lib.kmc_volume_min_max(ctypes.byref(v), None, ctypes.byref(mn), ctypes.byref(mx)) == 0
        v.has_min_max_values = 1
        v.min_value, v.max_value = mn.value, mx.value
        vols.append(t)
        specs.append((v, levels, kmc.POLYGON_NORMAL, True))
    torch.cuda.synchronize()
    nodes = sum(d[0] * d[1] * d[2] for d, _, _, _ in jobs)
    classified = sum(d[0] * d[1] * d[2] * len(lv) for d, _, _, lv in jobs)
    nsurf = sum(len(lv) for _, _, _, lv in jobs)

    out = kmc.Surface()

    def plain_all():
        tris = 0
        for v, levels, normal, dup in specs:
            for lv in levels:
                s = lib.kmc_extract(ctypes.byref(v), float(lv), normal, int(dup), kmc.MEMORY_DEVICE, None,
                                    ctypes.byref(out))
                if s != 0:
                    raise kmc.KmcError(s)
                tris += int(out.n_triangles)
                lib.kmc_surface_release(ctypes.byref(out))
        return tris

    print("[bench] batched leg: plain calls", file=sys.stderr, flush=True)
    plain_all()
    reps = max(1, args.batch_steps // 2)
    t0 = time.perf_counter()
    for _ in range(reps):
        tris_plain = plain_all()
    plain_ms = (time.perf_counter() - t0) / reps * 1e3

    print("[bench] batched leg: planning + autotuning the batch", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    b = kmc.Batch(specs, autotune=True)
    plan_s = time.perf_counter() - t0
    b.run()
    print(f"[bench] batched leg: timing {args.batch_steps} runs", file=sys.stderr, flush=True)
    dev_ms = []
    t0 = time.perf_counter()
    for _ in range(args.batch_steps):
        b.run()
        dev_ms.append(b.info()["last_run_ms"])
    batch_ms = (time.perf_counter() - t0) / args.batch_steps * 1e3
    info = b.info()
    b.close()
    assert info["n_triangles"] == tris_plain, (info["n_triangles"], tris_plain)
    return {
        "workload": f"{len(jobs)} float32 blob volumes, dims uniform 32..256 per axis, 1..16 blobs, "
                    f"1..4 isolevels in [0.2, 0.8] ({nsurf} surfaces), PolygonNormal + duplication",
        "volumes": len(jobs), "surfaces": nsurf, "nodes": nodes, "triangles": tris_plain,
        "plain_ms": round(plain_ms, 3),
        "batched_ms": round(batch_ms, 3),
        "batched_device_ms": round(float(np.median(dev_ms)), 3),
        "speedup": round(plain_ms / batch_ms, 2),
        "batched_gvoxels_per_s": round(classified / (batch_ms / 1e3) / 1e9, 2),
        "batched_mtriangles_per_s": round(tris_plain / (batch_ms / 1e3) / 1e6, 2),
        "plain_gvoxels_per_s": round(classified / (plain_ms / 1e3) / 1e9, 2),
        "plan_s": round(plan_s, 2),
        "autotuned": {"emit": "flat" if info["flat_emit"] else "tile teams", "tile_rows": info["tile_rows"],
                      "queue_capacity": info["queue_capacity"]},
        "graph_nodes": info["graph_nodes"], "passes": info["n_passes"],
    }


def serving_leg(lib, kmc, datasets, torch, args):
    """Serving on resident volumes: requests (volume, isolevel in the volume's
    10th..90th value percentile, PolygonNormal+duplication or VertexNormal)
    through kmc_server (continuous batching, 64 GiB budget) -- at saturation
    (all requests queued at once) and under Poisson arrivals at fractions of
    the saturation rate -- against the same requests as sequential plain
    kmc_extract calls.  Outputs stay on the device in both."""
    import numpy as np

    from kvs_amd.datasets import SplitMix64

    specs = datasets.serving_volumes(args.serve_volumes)
    print(f"[bench] serving leg: generating {len(specs)} volumes on the GPU", file=sys.stderr, flush=True)
    tensors, vols, pct = [], [], []
    nbytes = 0
    for edge, kind, seed in specs:
        if kind == "float32":
            t = datasets.turbulence_torch(edge, seed=seed, device="cuda")
        else:
            t = datasets.ct_like_torch(edge, seed=seed, dtype=np.dtype(kind), device="cuda")
        flat = t.reshape(-1)
        step = max(1, flat.numel() // 4_000_000)
        q = torch.quantile(flat[::step].to(torch.float64), torch.tensor([0.1, 0.9], dtype=torch.float64, device="cuda"))
        pct.append((float(q[0]), float(q[1])))
        tensors.append(t)
        vols.append(kmc.make_volume(t.data_ptr(), dims=(edge, edge, edge), memory=kmc.MEMORY_DEVICE, dtype=np.dtype(kind)))
        nbytes += t.numel() * t.element_size()
    torch.cuda.synchronize()
    server = kmc.Server(budget_bytes=64 << 30, workers=args.serve_workers, max_batch=args.serve_max_batch,
                        out_memory=kmc.MEMORY_DEVICE)
    vids = [server.register(v) for v in vols]
    # min/max as registered (the plain calls below use them too)
    for v in vols:
        mn, mx = ctypes.c_double(), ctypes.c_double()
        assert lib.kmc_volume_min_max(ctypes.byref(v), None, ctypes.byref(mn), ctypes.byref(mx)) == 0
        v.has_min_max_values, v.min_value, v.max_value = 1, mn.value, mx.value

    rng = SplitMix64(4100)
    reqs = []
    for _ in range(args.serve_requests):
        j = rng.integer(0, len(vols) - 1)
        lo, hi = pct[j]
        normal = kmc.POLYGON_NORMAL if rng.uniform() < 0.5 else kmc.VERTEX_NORMAL
        reqs.append((j, rng.uniform(lo, hi), normal, normal == kmc.POLYGON_NORMAL))

    # plain: one kmc_extract per request, in sequence
    print("[bench] serving leg: plain calls", file=sys.stderr, flush=True)
    out = kmc.Surface()
    t0 = time.perf_counter()
    tris_plain = 0
    for j, lv, normal, dup in reqs:
        st = lib.kmc_extract(ctypes.byref(vols[j]), float(lv), normal, int(dup), kmc.MEMORY_DEVICE, None,
                             ctypes.byref(out))
        if st != 0:
            raise kmc.KmcError(st)
        tris_plain += int(out.n_triangles)
        lib.kmc_surface_release(ctypes.byref(out))
    plain_s = time.perf_counter() - t0

    def serve(rate):
        """Submit every request (Poisson arrivals at `rate` per second, or all
        at once when rate is None), collect them; returns the server stats."""
        server.reset_stats()
        ids = [None] * len(reqs)
        arr = SplitMix64(4200)

        def producer():
            t_next = time.perf_counter()
            for k, (j, lv, normal, dup) in enumerate(reqs):
                if rate:
                    t_next += -math.log(1.0 - arr.uniform()) / rate
                    d = t_next - time.perf_counter()
                    if d > 0:
                        time.sleep(d)
                ids[k] = server.submit(vids[j], lv, normal, dup)

        th = threading.Thread(target=producer)
        th.start()
        th.join()
        tris = sum(server.wait_count(i) for i in ids)
        st = server.stats()
        assert tris == tris_plain and st["completed"] == len(reqs), (tris, tris_plain, st)
        return st

    print("[bench] serving leg: saturation", file=sys.stderr, flush=True)
    serve(None)   # warm
    sat = serve(None)
    sat_rate = len(reqs) / sat["span_s"]
    sweep = []
    for frac in (0.25, 0.5, 0.75, 0.9):
        print(f"[bench] serving leg: Poisson at {frac:.2f} of saturation", file=sys.stderr, flush=True)
        st = serve(frac * sat_rate)
        sweep.append({"offered_qps": round(frac * sat_rate, 1), "achieved_qps": round(len(reqs) / st["span_s"], 1),
                      "p50_ms": round(st["p50_ms"], 3), "p99_ms": round(st["p99_ms"], 3),
                      "mtriangles_per_s": round(st["triangles_per_s"] / 1e6, 1),
                      "peak_inflight": st["peak_inflight"], "batches": st["batches"], "deferred": st["deferred"]})
    server.close()
    return {
        "workload": f"{len(specs)} resident volumes (cubic, edge 64..512; float32 turbulence, uint16/uint8 "
                    f"CT-like; {nbytes / 2**30:.2f} GiB), {len(reqs)} requests: isolevel uniform in the "
                    f"volume's 10th..90th value percentile, PolygonNormal+duplication or VertexNormal "
                    f"(1:1); HBM budget 64 GiB; device-resident outputs",
        "requests": len(reqs), "triangles": tris_plain,
        "plain_ms_per_request": round(plain_s / len(reqs) * 1e3, 3),
        "plain_mtriangles_per_s": round(tris_plain / plain_s / 1e6, 1),
        "saturation": {"qps": round(sat_rate, 1), "mtriangles_per_s": round(sat["triangles_per_s"] / 1e6, 1),
                       "p50_ms": round(sat["p50_ms"], 3), "p99_ms": round(sat["p99_ms"], 3),
                       "batches": sat["batches"], "peak_inflight": sat["peak_inflight"]},
        "speedup_vs_plain": round(plain_s / sat["span_s"], 2),
        "poisson_sweep": sweep,
    }


def ct_leg(lib, kmc, datasets, torch, args):
    """512^3 uint8 CT-like volume (value noise + 20 ellipsoids), isolevel 100,
    VertexNormal: one kmc_extract on the HBM-resident volume (device outputs),
    and the PCIe-inclusive drop-in case (host volume in, host surface out);
    the reference's own kvs::MarchingCubes on the same bytes, 1 thread, its
    coordinates, connections and normals compared with ours bit for bit."""
    import numpy as np

    import oracle

    n, iso = 512, 100.0
    print("[bench] CT leg: generating 512^3 uint8 volume", file=sys.stderr, flush=True)
    t = datasets.ct_like_torch(n, seed=1001, nellipsoids=20, dtype=np.uint8, device="cuda")
    torch.cuda.synchronize()
    v = kmc.make_volume(t.data_ptr(), dims=(n, n, n), memory=kmc.MEMORY_DEVICE, dtype=np.uint8)
    mn, mx = ctypes.c_double(), ctypes.c_double()
    assert lib.kmc_volume_min_max(ctypes.byref(v), None, ctypes.byref(mn), ctypes.byref(mx)) == 0
    v.has_min_max_values, v.min_value, v.max_value = 1, mn.value, mx.value
    out = kmc.Surface()

    def once(vol, memory):
        st = lib.kmc_extract(ctypes.byref(vol), iso, kmc.VERTEX_NORMAL, 0, memory, None, ctypes.byref(out))
        if st != 0:
            raise kmc.KmcError(st)
        r = (int(out.n_vertices), int(out.n_triangles))
        lib.kmc_surface_release(ctypes.byref(out))
        return r

    once(v, kmc.MEMORY_DEVICE)
    reps = 5
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        nv, nt = once(v, kmc.MEMORY_DEVICE)
    dev_ms = (time.perf_counter() - t0) / reps * 1e3
    host = t.cpu().numpy()
    hv = kmc.make_volume(host, min_max=(mn.value, mx.value))
    once(hv, kmc.MEMORY_HOST)
    t0 = time.perf_counter()
    for _ in range(2):
        once(hv, kmc.MEMORY_HOST)
    e2e_ms = (time.perf_counter() - t0) / 2 * 1e3
    cpu = None
    if not args.no_cpu_baseline and oracle.ref_available():
        print("[bench] CT leg: reference CPU path", file=sys.stderr, flush=True)
        r = oracle.ref_marching_cubes(host, iso, oracle.VERTEX_NORMAL, False, min_max_values=(mn.value, mx.value))
        assert r["connections"].size // 3 == nt, (r["connections"].size // 3, nt)
        mine = kmc.extract(host, iso, kmc.VERTEX_NORMAL, False, min_max=(mn.value, mx.value))
        same = all(np.array_equal(np.ascontiguousarray(mine[k]).view(np.uint32),
                                  np.ascontiguousarray(r[k]).view(np.uint32)) for k in ("coords", "normals")) and \
            np.array_equal(mine["connections"], r["connections"])
        cpu = {"ms": round(r["seconds"] * 1e3, 1), "cores": 1, "kind": "reference",
               "speedup_device": round(r["seconds"] * 1e3 / dev_ms, 1),
               "gpu_output_identical": bool(same)}
        del mine, r
    return {
        "workload": "512^3 uint8 CT-like (32^3 value-noise lattice + 20 ellipsoids), isolevel 100, VertexNormal",
        "vertices": nv, "triangles": nt,
        "device_ms": round(dev_ms, 3),
        "gvoxels_per_s": round(n ** 3 / (dev_ms / 1e3) / 1e9, 2),
        "mtriangles_per_s": round(nt / (dev_ms / 1e3) / 1e6, 1),
        "end_to_end_ms": round(e2e_ms, 2),
        "end_to_end_note": "host volume in, host (pinned) surface out: PCIe both ways, drop-in case",
        "cpu_reference": cpu,
    }


def fp64_reevaluation(d32, iso, coords, lo, hi):
    """Largest deviation of shared-vertex surface coordinates from an FP64
    re-evaluation of the interpolation on the volume d32 ([z][y][x], unit
    spacing, clamp range [lo, hi]): the crossing edges are enumerated in node
    scan order x (x, y, z) edges, the order of the returned vertices, and
    interpolated in float64 (the same clamp and ratio, no rounding to float
    until the end).  The coordinates are the reference's bit for bit; this
    measures how far that float rounding is from exact arithmetic.  Returns
    (vertices, max_abs, max_rel)."""
    import numpy as np

    coords = coords.reshape(-1, 3)
    nz, ny, nx = d32.shape
    above = d32.astype(np.float64) > iso
    keys, pts = [], []
    for axis, (dz, dy, dx) in enumerate(((0, 0, 1), (0, 1, 0), (1, 0, 0))):
        cross = above[:nz - dz, :ny - dy, :nx - dx] != above[dz:, dy:, dx:]
        zz, yy, xx = np.nonzero(cross)
        del cross
        va = np.clip(d32[zz, yy, xx].astype(np.float64), lo, hi)
        vb = np.clip(d32[zz + dz, yy + dy, xx + dx].astype(np.float64), lo, hi)
        r = np.abs((iso - va) / (vb - va))
        pts.append(np.stack([xx + r * dx, yy + r * dy, zz + r * dz], axis=1))
        keys.append(((zz.astype(np.int64) * ny + yy) * nx + xx) * 3 + axis)
    del above
    order = np.argsort(np.concatenate(keys), kind="stable")
    exact = np.concatenate(pts)[order]
    assert exact.shape == coords.shape, (exact.shape, coords.shape)
    err = np.abs(coords.astype(np.float64) - exact)
    rel = err / np.maximum(np.abs(exact), 1e-30)
    big = np.abs(exact) > 1e-6
    return int(coords.shape[0]), float(err.max(initial=0.0)), float(rel[big].max(initial=0.0))


def rows_leg(lib, kmc, datasets, torch, vol, v, mn, mx, sptr, with_cpu):
    """The paths beside the isosurface, each on the GPU with the reference's
    own code timed on the host beside it (1 thread) and the outputs compared:

      slice_plane    kvs::SlicePlane through the centre of the resident 1024^3
                     volume (oblique plane, normal (1, 2, 3)); reference on a
                     32-slice z-slab cut by the same plane;
      unstructured   kvs::Isosurface over a tetrahedral UnstructuredVolumeObject
                     (6 tetrahedra per cube of a jittered 81^3 node grid, 3.07 M
                     cells, blob field, isolevel 0.5, PolygonNormal +
                     duplication), device-resident mesh and outputs;
      kvsml_ingest   a 1024 x 1024 x 256 float32 slab of the volume written as
                     KVSML with an external binary array (1 GiB, no min/max in
                     the header): read file -> HBM with the device min/max, vs
                     the reference's StructuredVolumeObject::read into host
                     memory; the file is in the page cache (written just before);
      stl_export     the slab's isolevel-0 VertexNormal surface written as a
                     binary STL: kmc_export_stl (device expansion) vs
                     PolygonExporter<Stl> + Stl::write; the files are compared
                     byte for byte; plus the device-to-device expansion alone;
      kvsml_export   the same surface as a KVSML polygon object, external
                     binary and ascii: kmc_export_kvsml vs PolygonExporter<
                     KVSMLPolygonObject> + write; every file compared."""
    import shutil
    import tempfile

    import numpy as np

    import oracle

    ref = with_cpu and oracle.ref_available()
    res = {}
    n = int(vol.shape[0])

    def timed(fn, reps):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            r = fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps * 1e3, r

    # SlicePlane
    print("[bench] rows leg: SlicePlane", file=sys.stderr, flush=True)
    table = kmc.colormap_table(0)
    tb = table.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
    nrm = np.array([1.0, 2.0, 3.0]) / np.sqrt(14.0)
    centre = (n - 1) / 2.0
    coef = [float(x) for x in np.float32(list(nrm) + [-nrm.sum() * centre])]
    cf = (ctypes.c_float * 4)(*coef)
    out = kmc.Surface()

    def slice_once():
        st = lib.kmc_slice_plane(ctypes.byref(v), cf, tb, 256, None, kmc.MEMORY_DEVICE, sptr, ctypes.byref(out))
        if st != 0:
            raise kmc.KmcError(st)
        t = int(out.n_triangles)
        lib.kmc_surface_release(ctypes.byref(out))
        return t

    ms, ntri = timed(slice_once, 10)
    sp = {"workload": f"{n}^3 float32, plane through the centre with normal (1,2,3)/sqrt(14), device outputs",
          "ms": round(ms, 3), "triangles": ntri, "gcells_per_s": round((n - 1) ** 3 / (ms / 1e3) / 1e9, 1),
          "mtriangles_per_s": round(ntri / (ms / 1e3) / 1e6, 1)}
    if ref:
        z0, k = n // 2 - 16, 32
        slab = vol[z0:z0 + k].cpu().numpy()
        c2 = [coef[0], coef[1], coef[2], float(np.float32(coef[3] + coef[2] * z0))]
        r = oracle.ref_slice_plane(slab, c2, tf_kind=0, min_max_values=(mn, mx))
        cells = (n - 1) * (n - 1) * (k - 1)
        sp["cpu_reference"] = {"sample": f"z-slab {n}x{n}x{k}, the same plane", "s": round(r["seconds"], 3),
                               "gcells_per_s": round(cells / r["seconds"] / 1e9, 4), "cores": 1}
        del slab, r
    res["slice_plane"] = sp

    # unstructured tetrahedra
    print("[bench] rows leg: unstructured tetrahedra (generating the mesh)", file=sys.stderr, flush=True)
    uv, uc, uk = datasets.unstructured_mesh((81, 81, 81), datasets.TETRAHEDRA, seed=11)
    tv, tc, tk = (torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (uv, uc, uk))
    torch.cuda.synchronize()
    u = kmc.UnstructuredVolume()
    u.values, u.coords, u.connections = tv.data_ptr(), tc.data_ptr(), tk.data_ptr()
    u.memory, u.type = kmc.MEMORY_DEVICE, kmc.NUMPY_TYPES[np.dtype(np.float32)]
    u.n_nodes, u.n_cells, u.cell_type, u.veclen = uv.size, uk.size // 4, datasets.TETRAHEDRA, 1

    def tet_once(keep=False):
        st = lib.kmc_extract_unstructured(ctypes.byref(u), 0.5, kmc.POLYGON_NORMAL, 1, kmc.MEMORY_DEVICE, sptr,
                                          ctypes.byref(out))
        if st != 0:
            raise kmc.KmcError(st)
        t = int(out.n_triangles)
        r = kmc.device_surface_to_numpy(out) if keep else None
        lib.kmc_surface_release(ctypes.byref(out))
        return t, r

    ms, (ntri, _) = timed(tet_once, 10)
    cells = uk.size // 4
    un = {"workload": f"{cells} tetrahedra (81^3 jittered node grid, 6 per cube), float32 blob field, isolevel 0.5, "
                      f"PolygonNormal + duplication, device-resident mesh and outputs",
          "ms": round(ms, 3), "triangles": ntri, "mcells_per_s": round(cells / (ms / 1e3) / 1e6, 1),
          "mtriangles_per_s": round(ntri / (ms / 1e3) / 1e6, 1)}
    if ref:
        _, mine = tet_once(keep=True)
        r = oracle.ref_unstructured_isosurface(uv, uc, uk, datasets.TETRAHEDRA, 0.5)
        same = all(np.asarray(mine[key]).tobytes() == np.asarray(r[key]).tobytes() for key in ("coords", "normals"))
        un["cpu_reference"] = {"s": round(r["seconds"], 4), "mcells_per_s": round(cells / r["seconds"] / 1e6, 2),
                               "cores": 1, "gpu_output_identical": bool(same)}
        del r, mine
    res["unstructured"] = un
    del tv, tc, tk, uv, uc, uk

    # KVSML ingestion and STL export
    tmp = tempfile.mkdtemp(prefix="kmc_bench_")
    try:
        print("[bench] rows leg: writing a 1 GiB KVSML file", file=sys.stderr, flush=True)
        k = 256
        slab = vol[:k].cpu().numpy()
        dat = os.path.join(tmp, "slab_value.dat")
        slab.tofile(dat)
        path = os.path.join(tmp, "slab.kvsml")
        with open(path, "w") as f:
            f.write('<?xml version="1.0" ?>\n<KVSML>\n    <Object type="StructuredVolumeObject">\n'
                    f'        <StructuredVolumeObject grid_type="uniform" resolution="{n} {n} {k}">\n'
                    '            <Node>\n                <Value veclen="1">\n'
                    '                    <DataArray type="float" format="binary" file="slab_value.dat" '
                    'endian="little" />\n                </Value>\n            </Node>\n'
                    '        </StructuredVolumeObject>\n    </Object>\n</KVSML>\n')
        print("[bench] rows leg: KVSML to HBM", file=sys.stderr, flush=True)
        kv = kmc.KvsmlVolume(path, kmc.MEMORY_DEVICE)
        kv.close()
        t0 = time.perf_counter()
        reps = 3
        for _ in range(reps):
            kv = kmc.KvsmlVolume(path, kmc.MEMORY_DEVICE)
            kv.close()
        ms = (time.perf_counter() - t0) / reps * 1e3
        kv = kmc.KvsmlVolume(path, kmc.MEMORY_DEVICE)
        same = bool(np.array_equal(kv.values(), slab)) and (kv.volume.min_value, kv.volume.max_value) == \
            (float(slab.min()), float(slab.max()))
        ki = {"workload": f"{n}x{n}x{k} float32 KVSML, external binary array ({slab.nbytes / 2**30:.2f} GiB), "
                          f"min/max computed at load; file in the page cache",
              "ms": round(ms, 2), "gbs": round(slab.nbytes / (ms / 1e3) / 1e9, 2), "values_identical": same}
        if ref:
            r = None
            t0 = time.perf_counter()
            r = oracle.ref_read_kvsml_volume(path)
            rs = time.perf_counter() - t0
            ki["cpu_reference"] = {"s": round(rs, 3), "gbs": round(slab.nbytes / rs / 1e9, 2), "cores": 1,
                                   "note": "StructuredVolumeObject::read into host memory, then a copy out "
                                           "of the reference's arrays (bound in)",
                                   "values_identical": bool(np.array_equal(r["values"], slab))}
            del r
        res["kvsml_ingest"] = ki

        print("[bench] rows leg: STL export", file=sys.stderr, flush=True)
        dev = kv.volume
        st = lib.kmc_extract(ctypes.byref(dev), 0.0, kmc.VERTEX_NORMAL, 0, kmc.MEMORY_DEVICE, sptr, ctypes.byref(out))
        if st != 0:
            raise kmc.KmcError(st)
        soup = kmc.Surface()

        def expand_once():
            e = lib.kmc_surface_expand(ctypes.byref(out), kmc.MEMORY_DEVICE, sptr, ctypes.byref(soup))
            if e != 0:
                raise kmc.KmcError(e)
            lib.kmc_surface_release(ctypes.byref(soup))

        ems, _ = timed(expand_once, 10)
        nt = int(out.n_triangles)
        host = kmc.device_surface_to_numpy(out)
        lib.kmc_surface_release(ctypes.byref(out))
        kv.close()
        mine = os.path.join(tmp, "mine.stl")
        t0 = time.perf_counter()
        kmc.export_stl(host, kmc.VERTEX_NORMAL, mine)
        gs = time.perf_counter() - t0
        ex = {"workload": f"isolevel 0 of the slab, VertexNormal: {nt} triangles, binary STL "
                          f"({os.path.getsize(mine) / 2**20:.0f} MiB)",
              "export_s": round(gs, 3),
              "expand_device_ms": round(ems, 3),
              "expand_gbs": round(nt * (12 + 72 + 36 + 12) / (ems / 1e3) / 1e9, 1),
              "expand_bytes_note": "per triangle: connections 12 B, 3 corners' coordinates and normals 72 B "
                                   "gathered, 36 + 12 B written"}
        if ref:
            theirs = os.path.join(tmp, "ref.stl")
            t0 = time.perf_counter()
            oracle.ref_export_stl(host, oracle.VERTEX_NORMAL, theirs)
            rs = time.perf_counter() - t0
            with open(mine, "rb") as a, open(theirs, "rb") as b:
                same = a.read() == b.read()
            ex["cpu_reference"] = {"s": round(rs, 3), "cores": 1, "file_identical": bool(same)}
        res["stl_export"] = ex

        print("[bench] rows leg: KVSML polygon export", file=sys.stderr, flush=True)
        color = (255, 255, 255)
        kx = {"workload": f"the same surface ({nt} triangles, VertexNormal, one colour and opacity) as a KVSML "
                          f"polygon object"}
        mesh = dict(host, colors=np.array(color, np.uint8), opacities=np.array([255], np.uint8), color_type=1,
                    normal_type=kmc.VERTEX_NORMAL)
        for writing, label in ((kmc.KVSML_EXTERNAL_BINARY, "external_binary"), (kmc.KVSML_ASCII, "ascii")):
            d_mine, d_ref = os.path.join(tmp, "mine_" + label), os.path.join(tmp, "ref_" + label)
            os.makedirs(d_mine)
            os.makedirs(d_ref)
            t0 = time.perf_counter()
            kmc.export_kvsml(host, kmc.VERTEX_NORMAL, os.path.join(d_mine, "s.kvsml"), writing, color=color)
            leg = {"s": round(time.perf_counter() - t0, 3),
                   "bytes": sum(os.path.getsize(os.path.join(d_mine, f)) for f in os.listdir(d_mine))}
            if ref:
                t0 = time.perf_counter()
                oracle.ref_export_kvsml_polygon(mesh, os.path.join(d_ref, "s.kvsml"), writing)
                rs = time.perf_counter() - t0
                same = sorted(os.listdir(d_mine)) == sorted(os.listdir(d_ref)) and all(
                    open(os.path.join(d_mine, f), "rb").read() == open(os.path.join(d_ref, f), "rb").read()
                    for f in os.listdir(d_mine))
                leg["cpu_reference"] = {"s": round(rs, 3), "cores": 1, "files_identical": bool(same)}
            kx[label] = leg
        res["kvsml_export"] = kx
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    return res


def cpu_baseline(vol, slices, vmin, vmax, kmc):
    """The reference's own kvs::MarchingCubes (oracle/_ref, 1 thread) on a
    z-slab sample of the same volume bytes, min/max pre-cached; the GPU's
    surfaces of the same slab (one kmc_extract_levels call, host in / host
    out) are compared with it bit for bit and re-evaluated in FP64.
    Returns (cpu_baseline, fp64_reevaluation)."""
    import numpy as np

    import oracle

    if not oracle.ref_available():
        return None, None
    n = vol.shape[2]
    z0 = (vol.shape[0] - slices) // 2
    sample = vol[z0:z0 + slices].cpu().numpy()
    nodes = sample.size
    gpu = kmc.extract_levels(sample, LEVELS, kmc.VERTEX_NORMAL, False, min_max=(vmin, vmax))
    total = 0.0
    tris = 0
    identical = True
    worst = {"vertices": 0, "max_abs": 0.0, "max_rel": 0.0}
    for lv, g in zip(LEVELS, gpu):
        print(f"[bench] cpu baseline level {lv}", file=sys.stderr, flush=True)
        r = oracle.ref_marching_cubes(sample, lv, oracle.VERTEX_NORMAL, False, min_max_values=(vmin, vmax))
        total += r["seconds"]
        tris += r["connections"].size // 3
        for k in ("coords", "connections", "normals"):
            a, b = np.asarray(g[k]), np.asarray(r[k])
            identical &= a.shape == b.shape and a.dtype == b.dtype and a.tobytes() == b.tobytes()
        del r
        nv, ea, er = fp64_reevaluation(sample, lv, g["coords"], vmin, vmax)
        worst["vertices"] += nv
        worst["max_abs"] = max(worst["max_abs"], ea)
        worst["max_rel"] = max(worst["max_rel"], er)
    cpu = {
        "value": round(nodes * len(LEVELS) / total / 1e9, 4),
        "unit": "Gvoxels/s",
        "cores": 1,
        "kind": "reference",
        "sample": f"z-slab {n}x{n}x{slices} of the same volume ({nodes} nodes), {len(LEVELS)} levels, "
                  f"VertexNormal, kvs::MarchingCubes constructor wall time {total:.2f}s, {tris} triangles",
        "mtriangles_per_s": round(tris / total / 1e6, 3),
        "gpu_output_identical": bool(identical),
        "cpu_model": cpu_model(),
        "nproc": os.cpu_count(),
    }
    fp64 = dict(worst, workload=f"the same z-slab, {len(LEVELS)} levels, shared vertices: coordinates against "
                                f"an FP64 evaluation of the same interpolation")
    return cpu, fp64


if __name__ == "__main__":
    main()
