#!/usr/bin/env python3
"""Benchmark: closest-point (UDF) queries/s on one MI355X at 4.1M triangles.

Workload (the reference's headline configuration): the union of 200
icospheres subdivided 5 times (4,096,000 faces, seed 1) and 2^24 query points
uniform in [-1, 1]^3.  One step = one closest-point pass over all 2^24
queries (distance, closest face, closest point and barycentrics), inputs
resident in HBM.  The BVH is built once before timing (build time reported
separately).

Prints one JSON line.  ``roofline`` is computed for the traversal kernel from
its HIP-event-timed average launch time and the algorithmic bytes per query
of the reference algorithm, B_q = 44 + 32*N_node + 36*N_tri, where N_node and
N_tri are the node visits and triangle tests per query of the reference's SAH
tree walk, counted by the CPU oracle on a sample of the same queries.
``cpu_baseline`` times that oracle walk (the reference algorithm restated in
C++) on this host's cores on a bounded sample.  ``surrogate_fallback``
summarises the learned surrogate + exact-BVH-fallback router on the same
queries (speed, error against the BVH, whether it meets the 1e-5 x bbox
diagonal max-error bar); ``value`` is the plain BVH's unless it does.

Other workloads (``--workload``): dmc, octree, flood, surrogate.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
N_QUERIES = 1 << 24


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--queries", type=int, default=N_QUERIES)
    ap.add_argument("--cpu-sample", type=int, default=1 << 21, help="queries in the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-surrogate", action="store_true", help="skip the surrogate+fallback measurement")
    ap.add_argument("--workload", choices=["udf", "dmc", "octree", "flood", "surrogate", "batch", "clip", "c1", "ray"],
                    default="udf",
                    help="udf: the closest-point headline; dmc: SparseDiffDMC forward+backward on a 1024^3 band; "
                         "octree: SAT voxelisation levels 4-10 of the 4.1M-face mesh; flood: sparse flood fill "
                         "of its 1024^3 voxels")
    ap.add_argument("--conn", type=int, default=26, help="flood connectivity (6, 18 or 26)")
    ap.add_argument("--dmc-res", type=int, default=1024)
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="before timing, run one plain device copy of this many MiB and the VALU calibration kernel "
                         "(counter calibration)")
    return ap.parse_args()


def profile_figures(name):
    """``key: value`` figures of a committed rocprofv3 summary
    (profiles/<name>.md), or {} when absent."""
    p = os.path.join(ROOT, "profiles", name + ".md")
    if not os.path.exists(p):
        return {}
    import re

    out = {}
    for m in re.finditer(r"^([a-z_0-9]+):\s*([-0-9.eE+]+)\s*$", open(p).read(), re.M):
        out[m.group(1)] = float(m.group(2))
    return out


def roofline(kernel, kernel_ms, profile, algorithmic=None):
    """The roofline object of a kernel from its committed counter summary:
    achieved = the HBM bytes it moves per launch (FETCH_SIZE + WRITE_SIZE,
    scaled by the copy-kernel calibration) / its HIP-event-timed launch time,
    frac = achieved / 8 TB/s; plus its VALU busy fraction (rocprofv3's
    VALUBusy), VALU instructions per wave, L2 hit rate and occupancy where
    collected.  ``algorithmic`` (bytes per launch of the
    reference algorithm's work, and a note) is reported beside it, labelled:
    it counts work this kernel skips, so it is not a rate it moves."""
    f = profile_figures(profile)
    sec = kernel_ms * 1e-3
    traffic = f.get("hbm_bytes_per_launch")
    r = {"bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": None, "traffic": traffic,
         "kernel": kernel, "kernel_ms": kernel_ms, "profile": f"profiles/{profile}.md" if f else None}
    if traffic:
        r["achieved"] = traffic / sec / 1e9
        r["frac"] = r["achieved"] / HBM_PEAK_GBS
    if "valu_issue_per_cu_cycle" in f and "valu_issue_calib_per_cu_cycle" in f:
        # VALU issue: the kernel's SQ_INSTS_VALU per CU per cycle (GRBM_GUI_ACTIVE) against the rate the
        # calibration kernel (dependency-free FP32 fused multiply-adds) reaches with the same counters; a
        # kernel at or above it issues VALU instructions as fast as plain FP32 work does (its mix holds
        # instructions cheaper than an FMA), the bound of a kernel that moves little memory
        r["valu_issue"] = {"achieved": f["valu_issue_per_cu_cycle"], "calib": f["valu_issue_calib_per_cu_cycle"],
                           "unit": "wave-instructions per CU per cycle",
                           "ratio": f["valu_issue_per_cu_cycle"] / f["valu_issue_calib_per_cu_cycle"]}
    for k in ("valu_insts_per_wave", "salu_insts_per_wave", "l2_hit_rate", "occupancy_percent",
              "valu_utilization_percent", "valu_busy_percent", "kernel_ms_profiled"):
        if k in f:
            r[k] = f[k]
    if algorithmic:
        b, note = algorithmic
        r["algorithmic"] = {"bytes_per_launch": b, "GBps": b / sec / 1e9, "note": note}
    return r


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def median_of(fn, runs=5):
    """One warm-up run, then the median wall time of `runs` runs (s)."""
    fn()
    ts = []
    for _ in range(runs):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts))


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {a.gpus}); "
              "multi-GPU query sharding is not built.", file=sys.stderr)
        sys.exit(2)

    if a.copy_probe_mb:
        # counter calibration: one elementwise kernel reading and writing
        # copy_probe_mb MiB (the HBM byte counters' scale), and the VALU
        # calibration kernel (the issue rate of plain FP32 VALU instructions)
        import torch

        from atom3d_amd.kernels import _lib

        src = torch.ones(a.copy_probe_mb << 18, device="cuda:0")
        dst = torch.empty_like(src)
        torch.mul(src, 2.0, out=dst)
        torch.cuda.synchronize()
        del src, dst
        out = torch.empty(256 * 8192, device="cuda:0")
        _lib.call("a3d_calib_valu", _lib.ptr(out), out.numel(), 4096, _lib.stream(out.device))
        torch.cuda.synchronize()
        del out
    if a.workload == "dmc":
        return bench_dmc(a)
    if a.workload in ("octree", "flood"):
        return bench_octree_flood(a)
    if a.workload == "surrogate":
        return bench_surrogate(a)
    if a.workload == "batch":
        return bench_batch(a)
    if a.workload == "clip":
        return bench_clip(a)
    if a.workload == "ray":
        return bench_ray(a)
    if a.workload == "c1":
        return bench_c1(a)

    import torch

    from atom3d_amd import MeshBVH, datasets
    from atom3d_amd.kernels import _lib
    from atom3d_amd.kernels.bvh import BVHAccelerator

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    V, F = datasets.sphere_union()
    P = datasets.uniform_points(a.queries, seed=1).to(dev)
    Vd, Fd = V.to(dev), F.to(dev)
    K = P.shape[0]
    torch.cuda.synchronize()
    # runtime start-up (kernel loading, allocator pools) outside every timing
    BVHAccelerator(Vd[:3 * 64], Fd[:64] % 64).udf(P[:1 << 16])
    torch.cuda.synchronize()

    # cold call: a fresh tree and one closest-point call on it (Morton order +
    # the plain walk: a mesh's first call never builds the distance grid)
    t0 = time.perf_counter()
    bvh = BVHAccelerator(Vd, Fd)
    torch.cuda.synchronize()
    build_s = time.perf_counter() - t0
    bvh.udf(P)
    torch.cuda.synchronize()
    cold_s = time.perf_counter() - t0

    dist = torch.empty(K, device=dev)
    face = torch.empty(K, dtype=torch.int32, device=dev)
    closest = torch.empty(K, 3, device=dev)
    uvw = torch.empty(K, 3, device=dev)
    perm = torch.empty(K, dtype=torch.int32, device=dev)
    wsb = ctypes.c_int64()
    _lib.call("a3d_query_order_work_ws_bytes", K, ctypes.byref(wsb))
    ws = torch.empty(wsb.value, dtype=torch.uint8, device=dev)
    stream = torch.cuda.current_stream(dev)
    sp = stream.cuda_stream
    lib = _lib.lib()
    # the distance grid that bounds each walk: built once per mesh, like the
    # tree (what MeshBVH.udf does once a mesh has answered a few grids' worth
    # of queries); timed twice: the process's first build and a rebuild
    torch.cuda.synchronize()
    tg = time.perf_counter()
    bvh.seed_grid()
    torch.cuda.synchronize()
    grid_first_s = time.perf_counter() - tg
    bvh.__dict__.pop("_seed_grids")
    tg = time.perf_counter()
    grid, gface, glo, gh = bvh.seed_grid()
    torch.cuda.synchronize()
    grid_s = time.perf_counter() - tg
    work = bvh.__dict__["_seed_work"][bvh.SEED_GRID_RES]
    order_args = (_lib.ptr(P), K, _lib.ptr(work), bvh.SEED_GRID_RES, (ctypes.c_float * 3)(*glo), ctypes.c_float(gh),
                  _lib.ptr(perm), _lib.ptr(ws), wsb.value, sp)
    udf_args = (_lib.ptr(bvh.blob), _lib.ptr(P), K, _lib.ptr(perm), _lib.ptr(grid), _lib.ptr(gface),
                _lib.ptr(bvh.vertices), _lib.ptr(bvh.faces), bvh.SEED_GRID_RES, (ctypes.c_float * 3)(*glo),
                ctypes.c_float(gh), ctypes.c_float(gh / 1024), _lib.ptr(dist), _lib.ptr(face), _lib.ptr(closest),
                _lib.ptr(uvw), 0, sp)

    # one step = ordering of the queries (Morton packets, heaviest cost class
    # of the grid's walk work first) + the wave-coherent walk, each query's
    # walk bounded from its start by the grid (plus the walk of any query
    # whose bound fell short)
    def step(ev=None):
        _lib.check(lib.a3d_query_order_work(*order_args), "a3d_query_order_work")
        if ev is not None:
            ev[0].record(stream)
        _lib.check(lib.a3d_bvh_udf_grid(*udf_args), "a3d_bvh_udf_grid")
        if ev is not None:
            ev[1].record(stream)

    for _ in range(a.warmup):
        step()
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(a.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(a.steps):
        step(evs[k])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern_ms = sum(e0.elapsed_time(e1) for e0, e1 in evs) / a.steps  # the walk kernel alone
    ms_per_step = wall * 1e3 / a.steps
    qps = K / (ms_per_step * 1e-3)

    # --- oracle: reference work counts and CPU baseline on bounded samples
    cpu = None
    n_node = n_tri = None
    if not a.no_cpu:
        import oracle

        cores = len(os.sched_getaffinity(0))
        tb = time.perf_counter()
        sah = oracle.SahBvh(V.numpy(), F.numpy())
        sah_build_s = time.perf_counter() - tb
        n_all, n_one = min(1 << 20, K), min(1 << 15, K)
        sample = P[:n_all].cpu().numpy()
        t_all = median_of(lambda: sah.udf(sample, threads=cores))
        t_one = median_of(lambda: sah.udf(sample[:n_one], threads=1))
        _, _, _, _, nv, nt = sah.udf(sample, threads=cores, counts=True)
        n_node, n_tri = float(nv.mean()), float(nt.mean())
        cpu = {
            "value": n_all / t_all,
            "unit": "queries/s",
            "cores": cores,
            "kind": "port",
            "single_thread_value": n_one / t_one,
            "cpu_model": cpu_model(),
            "protocol": "one warm-up run, then the median of 5",
            "sample": f"all cores: the first {n_all} of the {K} queries; single thread: the first {n_one}; the "
                      f"reference's SAH tree walk restated in C++ (OpenMP), on the reference's SAH tree of the "
                      f"same 4.1M-face mesh, whose single-threaded host build took {sah_build_s:.1f} s (not "
                      f"included)",
            "sah_build_s": sah_build_s,
        }
    if n_node is None:
        n_node, n_tri = 245.06, 212.58  # oracle counts on a 2^20-query sample of this workload
    bq = 44.0 + 32.0 * n_node + 36.0 * n_tri
    sg = None
    if not a.no_surrogate:
        r = measure_surrogate(MeshBVH(Vd, Fd), V, P, a, fp64_check=False)
        sg = {"queries_per_s": r["queries_per_s"], "speedup_vs_plain_bvh": qps and r["queries_per_s"] / qps,
              "max_err": r["errors_vs_bvh"]["max_err"], "mean_err": r["errors_vs_bvh"]["mean_err"],
              "frac_exact": r["errors_vs_bvh"]["frac_exact"], "bbox_diag": r["bbox_diag"],
              "meets_max_err_1e-5_diag": r["meets_max_err_1e-5_diag"], "train_s": r["train"]["train_s"],
              "note": "value is the plain exact BVH: the surrogate+fallback router does not meet the "
                      "1e-5 x bbox-diagonal max-error bar (see max_err)"}
    out = {
        "metric": "closest-point queries/s at 4M tris; surrogate+fallback vs plain BVH at 1e-5",
        "value": qps,
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": "plain BVH UDF: union of 200 icospheres (sub 5, seed 1, 4,096,000 faces), "
                        f"{K} uniform queries in [-1,1]^3; outputs dist, face, closest point, uvw; each "
                        f"walk bounded by a {bvh.SEED_GRID_RES + 1}^3 grid of exact distances built once per "
                        "mesh, like the tree (exact results)",
            "faces": int(F.shape[0]),
            "queries": K,
            "parallelism": "single GPU",
        },
        "roofline": roofline(
            "k_udf_wave", kern_ms, "k_udf_wave_sphere_union_counters",
            (bq * K, f"reference algorithm: B_q = 44 + 32 N_node + 36 N_tri = {bq:.0f} B per query with the "
                     f"reference SAH walk's N_node = {n_node:.1f}, N_tri = {n_tri:.1f} (oracle counts)")),
        "cpu_baseline": cpu,
        "bvh_build_s": build_s,
        "cold_call": {"queries_per_s": K / cold_s, "s": cold_s,
                      "what": "a fresh BVHAccelerator on the mesh + one udf() call of all queries (Morton order "
                              "and the plain walk: no grid on a first call)"},
        "seed_grid_build_s": grid_s,
        "seed_grid_first_build_s": grid_first_s,
        "surrogate_fallback": sg,
    }
    print(json.dumps(out), flush=True)


def bench_dmc(a):
    """SparseDiffDMC (training mode, float32) on the narrow band of 32 spheres
    and 16 tori (seed 2) at res^3: one step = forward (topology + values) and
    backward of sum(W * outputs) to sdf, deform, alpha and beta."""
    import torch

    from atom3d_amd import datasets
    from atom3d_amd.mesh_extractor import SparseDiffDMC

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    res = a.dmc_res
    tg = time.perf_counter()
    vox, sdf, ci = datasets.dmc_band(res, device=dev)
    torch.cuda.synchronize()
    gen_s = time.perf_counter() - tg
    N, M = vox.shape[0], sdf.shape[0]
    g = torch.Generator(device=dev).manual_seed(2)
    deform = (torch.rand(M, 3, device=dev, generator=g) - 0.5) * (0.2 / res)
    alpha = torch.randn(N, 8, device=dev, generator=g)
    beta = torch.randn(N, 12, device=dev, generator=g)
    leaves = [t.requires_grad_() for t in (sdf, deform, alpha, beta)]
    m = SparseDiffDMC(device=dev)
    W = {}

    def fwd():
        return m(vox, leaves[0], ci, res, deform=leaves[1], alpha=leaves[2], beta=leaves[3])

    def step(ev=None):
        for t in leaves:
            t.grad = None
        if ev is not None:
            ev[0].record()
        o = fwd()
        if ev is not None:
            ev[1].record()
        loss = sum((W[k] * getattr(o, k)).sum() for k in W)
        loss.backward()
        if ev is not None:
            ev[2].record()
        return o

    o = fwd()
    for k in ("vertices", "L_dev", "p_alpha", "n_alpha"):
        W[k] = torch.randn(getattr(o, k).shape, device=dev, generator=g)
    V, Ftri, E = o.vertices.shape[0], o.faces.shape[0], o.L_dev.shape[0]
    del o
    for _ in range(a.warmup):
        step()
    torch.cuda.synchronize()
    evs = [tuple(torch.cuda.Event(enable_timing=True) for _ in range(3)) for _ in range(a.steps)]
    t0 = time.perf_counter()
    for k in range(a.steps):
        step(evs[k])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    fwd_ms = sum(e[0].elapsed_time(e[1]) for e in evs) / a.steps
    bwd_ms = sum(e[1].elapsed_time(e[2]) for e in evs) / a.steps
    ms = wall * 1e3 / a.steps
    n_surf = int(((sdf.detach()[ci.long()] < 0).sum(1) % 8 != 0).sum())
    fwd_bytes = 48 * N + 4 * M + 12 * V + 24 * Ftri
    bwd_bytes = 12 * V + 4 * M + 48 * n_surf

    # eval mode: the reference's default QEF placement (corner gradients
    # estimated, one regularised 10x3 least-squares solve per dual vertex)
    m.eval()
    det = [t.detach() for t in leaves]

    def qef():
        return m(vox, det[0], ci, res, deform=det[1])

    for _ in range(max(1, a.warmup)):
        qef()
    torch.cuda.synchronize()
    tq = time.perf_counter()
    for _ in range(a.steps):
        oq = qef()
    torch.cuda.synchronize()
    qef_ms = (time.perf_counter() - tq) * 1e3 / a.steps
    qef_out = {"ms": qef_ms, "cubes_per_s": N / (qef_ms * 1e-3), "vertices": int(oq.vertices.shape[0]),
               "faces": int(oq.faces.shape[0])}
    del oq
    m.train()

    cpu = None
    if not a.no_cpu:
        from oracle.dmc import dmc_forward

        rs = 128
        cv, cs, cc = datasets.dmc_band(rs, device="cpu")
        cpu_leaves = [cs.clone().requires_grad_(), torch.zeros(cs.shape[0], 3, requires_grad=True),
                      torch.randn(cv.shape[0], 8).requires_grad_(), torch.randn(cv.shape[0], 12).requires_grad_()]
        tc = time.perf_counter()
        oc = dmc_forward(cv.long(), cpu_leaves[0], cc.long(), rs, deform=cpu_leaves[1], alpha=cpu_leaves[2],
                         beta=cpu_leaves[3], dtype=torch.float32)
        sum(oc[k].sum() for k in ("vertices", "L_dev", "p_alpha", "n_alpha")).backward()
        cpu_s = time.perf_counter() - tc
        cpu = {"value": cv.shape[0] / cpu_s, "unit": "cubes/s", "cores": torch.get_num_threads(), "kind": "port",
               "sample": f"the same shapes' band at {rs}^3 ({cv.shape[0]} cubes), forward + backward of the "
                         f"reference's weighted-average algorithm restated in torch CPU ops (oracle/dmc.py), "
                         f"float32"}
        tc = time.perf_counter()
        dmc_forward(cv.long(), cs, cc.long(), rs, dtype=torch.float32, training=False, mode="qef")
        qef_out["cpu_port_cubes_per_s"] = cv.shape[0] / (time.perf_counter() - tc)
        qef_out["cpu_port_sample"] = f"eval-mode QEF forward of the {rs}^3 band ({cv.shape[0]} cubes), torch CPU"
    out = {
        "metric": "SparseDiffDMC cubes/s (forward + backward)",
        "value": N / (ms * 1e-3),
        "unit": "cubes/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": f"SparseDiffDMC training step on the |sdf(centre)| < h*sqrt(3)/2 band of 32 spheres "
                               f"+ 16 tori (seed 2) at {res}^3, deform/alpha/beta/sdf all requiring grad",
                   "cubes": N, "corners": M, "surface_cubes": n_surf, "vertices": V, "faces": Ftri,
                   "entries": E, "parallelism": "single GPU"},
        "forward_ms": fwd_ms,
        "backward_ms": bwd_ms,
        "forward_cubes_per_s": N / (fwd_ms * 1e-3),
        "qef_eval": qef_out,
        "roofline": roofline("whole step (forward + backward kernels)", ms, "dmc_band1024_counters",
                             (fwd_bytes + bwd_bytes,
                              f"forward 48 N + 4 M + 12 V + 24 F = {fwd_bytes} B, backward 12 V + 4 M + 48 N_surf "
                              f"= {bwd_bytes} B (inputs read and outputs written once)")),
        "cpu_baseline": cpu,
        "band_build_s": gen_s,
    }
    print(json.dumps(out), flush=True)


def bench_octree_flood(a):
    """octree: OctreeIndexer(max_level=10).octree_traverse(min_level=4) on the
    4.1M-face union of icospheres, one step = the whole level loop; flood: the
    sparse flood fill of its level-10 voxels (crop = voxel bounding box +- 10,
    source (0,0,0), k-ring 1), one step = fill + classification + extraction
    of every output list."""
    import torch

    from atom3d_amd import MeshBVH, OctreeIndexer, datasets
    from atom3d_amd.kernels.flood_fill import PLANE, flood_planes

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    V, F = datasets.sphere_union()
    mesh = MeshBVH(V.to(dev), F.to(dev))
    torch.cuda.synchronize()
    tb = time.perf_counter()
    _ = mesh._bvh.box_blob  # the box queries' tree (the reference's SAH tree, built on the device)
    torch.cuda.synchronize()
    sah_build_s = time.perf_counter() - tb
    oc = OctreeIndexer(max_level=10, device=dev)
    vox = oc.octree_traverse(mesh, min_level=4)
    torch.cuda.synchronize()
    counts = list(oc.last_level_counts)
    tested = sum(c[1] for c in counts)
    n_vox = int(vox.shape[0])
    names = ("water", "dry", "collision", "dam", "ext_tide", "dry_adj")

    if a.workload == "octree":
        def step():
            return oc.octree_traverse(mesh, min_level=4)
    else:
        def step():
            st = flood_planes(vox, 1024, (0, 0, 0), 10, a.conn, 1)
            return [st.coords(k) for k in names], st
    for _ in range(a.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        r = step()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / a.steps

    cpu = None
    n_node = n_tri = None
    if a.workload == "octree":
        if not a.no_cpu:
            import oracle

            cores = len(os.sched_getaffinity(0))
            tb = time.perf_counter()
            sah = oracle.SahBvh(V.numpy(), F.numpy())
            sah_s = time.perf_counter() - tb
            tq = time.perf_counter()
            ref, ref_tested, _, (wn, wt) = sah.octree([-1, -1, -1, 1, 1, 1], 4, 10, threads=cores, full=True,
                                                      work=True)
            cpu_s = time.perf_counter() - tq
            assert np.array_equal(ref, vox.cpu().numpy()), "octree differs from the oracle"
            n_node, n_tri = wn / ref_tested.sum(), wt / ref_tested.sum()
            cpu = {"value": n_vox / cpu_s, "unit": "voxels/s", "cores": cores, "kind": "port",
                   "sample": f"the whole workload (levels 4-10, {int(ref_tested.sum())} cube tests); the "
                             f"reference's level loop with its full SAT pair walk restated in C++ (OpenMP, "
                             f"{cores} threads) on the reference SAH tree; SAH build {sah_s:.1f} s not included"}
        if n_node is None:
            n_node, n_tri = 29.7, 4.4
        unit_bytes = 13 + 32 * n_node + 36 * n_tri
        units, kernel_name, kernel_profile = tested / len(counts), "k_octree_test", "k_octree_test_sphere_union_counters"
        alg_note = (f"reference algorithm per cube test: 13 + 32 N_node + 36 N_tri = {unit_bytes:.0f} B (its full pair "
                    f"walk, oracle counts), times the mean cube tests per level launch")
        value, unit = n_vox / (ms * 1e-3), "voxels/s"
        extra = {"cubes_tested": tested, "cube_tests_per_s": tested / (ms * 1e-3), "levels": counts,
                 "n_node": n_node, "n_tri": n_tri, "box_tree_build_s": sah_build_s,
                 "box_tree_depth": mesh._bvh.box_depth}
        metric = "octree SAT voxelisation voxels/s (levels 4-10, 4.1M tris)"
        workload = "OctreeIndexer(max_level=10).octree_traverse(min_level=4) on the union of 200 icospheres " \
                   "(4,096,000 faces, seed 1), bounds [-1,1]^3"
    else:
        lists, st = r
        crop = st.D * st.H * st.W
        f_water = st.totals[PLANE["water"]] / crop
        unit_bytes = 1 / 8 + 4 + 8 * f_water
        units, kernel_name, kernel_profile = crop, "whole step", "flood_sphere_union_1024_counters"
        alg_note = "per crop voxel: 1/8 B occupancy + 4 B label + 8 B per water voxel (the reference's BFS)"
        value, unit = crop / (ms * 1e-3), "crop voxels/s"
        extra = {"crop": [st.D, st.H, st.W], "passes": st.passes, "f_water": f_water,
                 "counts": {k: int(x.shape[0]) for k, x in zip(names, lists)},
                 "glue": glue_leg(vox, dict(zip(names, lists)), oc, a)}
        if not a.no_cpu:
            from oracle.flood import sparse_flood_fill_from_voxels

            sub = vox[(vox < 512).all(1)].cpu().numpy()
            tq = time.perf_counter()
            sparse_flood_fill_from_voxels(sub, 512, (0, 0, 0), a.conn, 1)
            cpu_s = time.perf_counter() - tq
            lo = np.maximum(sub.min(0) - 10, 0)
            hi = np.minimum(sub.max(0) + 10, 511)
            ccrop = int(np.prod(hi - lo + 1))
            cpu = {"value": ccrop / cpu_s, "unit": "crop voxels/s", "cores": 1, "kind": "port",
                   "sample": f"the voxels of the same mesh inside the [0,512)^3 octant ({len(sub)} dam voxels, "
                             f"{ccrop} crop voxels); the reference's BFS restated in C++ (single thread) plus the "
                             f"numpy restatement of its classification and k-ring signs"}
        metric = f"sparse flood fill crop voxels/s (1024^3, conn {a.conn})"
        workload = f"sparse_flood_fill of the union of 200 icospheres' level-10 voxels ({n_vox}), padding 10, " \
                   f"source (0,0,0), connectivity {a.conn}, k-ring 1"
    prof = profile_figures(kernel_profile)
    kms = prof.get("kernel_ms_profiled", ms)
    rl = roofline(kernel_name, kms, kernel_profile, (unit_bytes * units, alg_note))
    if "kernel_ms_profiled" not in prof:
        rl["kernel"] = "whole step (no kernel profile)"
    out = {
        "metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": workload, "voxels": n_vox, "parallelism": "single GPU", **extra},
        "roofline": rl,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def glue_leg(vox, lists, oc, a):
    """Timings of the formats between the flood fill / octree and the SDF /
    DMC steps at 1024^3: construct_surface_dual_graph of the flood's dam,
    collision and dry-adjacent voxels, the wet-corner mask of the dam, and
    voxel_unique_vertices of the octree's voxels (ms per call, best of
    steps), with the numpy oracle on the same inputs as a CPU reference."""
    import torch

    from atom3d_amd.apps import construct_surface_dual_graph, get_dam_wet_corners

    dam, coll, dry = (lists[k].int() for k in ("dam", "collision", "dry_adj"))
    flat = oc.ijk_to_cube(vox)

    def best(fn):
        fn()
        torch.cuda.synchronize()
        ts = []
        for _ in range(max(1, a.steps)):
            t0 = time.perf_counter()
            r = fn()
            torch.cuda.synchronize()
            ts.append((time.perf_counter() - t0) * 1e3)
        return min(ts), r

    dg_ms, g = best(lambda: construct_surface_dual_graph(dam, coll, dry, oc))
    wc_ms, _ = best(lambda: get_dam_wet_corners(dam, coll, 1024))
    uv_ms, u = best(lambda: oc.voxel_unique_vertices(flat))
    out = {"dual_graph_ms": dg_ms, "cells": int(g[4].shape[0]), "dual_vertices": int(g[1].shape[0]),
           "dual_graph_cells_per_s": int(g[4].shape[0]) / (dg_ms * 1e-3), "wet_corners_ms": wc_ms,
           "unique_vertices_ms": uv_ms, "unique_vertices_voxels": int(vox.shape[0]),
           "unique_vertices": int(u[0].shape[0]), "unique_vertices_voxels_per_s": int(vox.shape[0]) / (uv_ms * 1e-3)}
    if not a.no_cpu:
        from oracle import cells as ocl

        d, c, y = ((x[(x < 384).all(1)]).cpu().numpy() for x in (dam, coll, dry))
        t0 = time.perf_counter()
        r = ocl.construct_surface_dual_graph(d, c, y, 1024, [-1.0] * 3, [1.0] * 3)
        cpu_s = time.perf_counter() - t0
        out["dual_graph_cpu_port_cells_per_s"] = len(r[4]) / cpu_s
        out["cpu_port"] = (f"numpy restatement (oracle/cells.py), single thread, on the cells inside [0,384)^3 "
                           f"({len(r[4])} cells)")
    return out


SWEEP_BANDS = (0.002, 0.005, 0.01, 0.02, 0.05, 0.1, 0.2, 0.5, 4.0)


def measure_surrogate(mesh, V, P, a, fp64_check: bool, sweep: bool = False):
    """Train the surrogate on the mesh (2^24 BVH-labelled samples, half
    uniform in [-1,1]^3, half within 0.02 of the surface; 2000 Adam steps of
    2^18) and time the router on P: Morton ordering + network on every query
    + compaction + BVH on the queries predicted below the 0.01 band."""
    import torch

    from atom3d_amd.kernels import _lib
    from atom3d_amd.surrogate import UDFSurrogate

    dev = P.device
    K = P.shape[0]
    band = 0.01
    sg = UDFSurrogate(mesh, seed=0)
    hist = sg.fit(n_samples=1 << 24, steps=2000, batch=1 << 18, lr=1e-2, log_every=500)
    sp = torch.cuda.current_stream(dev)

    def timed(fn, n):
        for _ in range(a.warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(n):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3 / n

    ms = timed(lambda: sg.udf(P, band=band), a.steps)
    perm = mesh._bvh.query_order(P)
    out = torch.empty(K, device=dev)
    args = (_lib.ptr(sg.params), _lib.ptr(P), K, _lib.ptr(perm), _lib.ptr(out), sp.cuda_stream)
    for _ in range(2):
        _lib.call("a3d_sg_infer", *args)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(sp)
    for _ in range(a.steps):
        _lib.call("a3d_sg_infer", *args)
    e1.record(sp)
    torch.cuda.synchronize()
    infer_ms = e0.elapsed_time(e1) / a.steps
    ev = sg.evaluate(P, band=band)
    diag = float((V.max(0)[0] - V.min(0)[0]).norm())
    flops = 2 * (32 * 64 + 64 * 64 + 64)
    tflops = flops * K / (infer_ms * 1e-3) / 1e12
    fp64 = None
    if fp64_check:
        import oracle

        F = mesh.faces.cpu()
        sub = P[:256].cpu().numpy()
        d64, _ = oracle.fp64_udf(V.numpy(), F.numpy(), sub)
        dr = sg.udf(P[:256], band=band)[0].cpu().numpy().astype(np.float64)
        fp64 = {"n": 256, "max_err": float(np.abs(dr - d64).max()), "mean_err": float(np.abs(dr - d64).mean())}
    rows = []
    if sweep:  # the safety band against speed and error (4.0 > the box diagonal: every query exact)
        ref = mesh._bvh.udf(P)[0]
        for b in SWEEP_BANDS:
            bms = timed(lambda: sg.udf(P, band=b), max(3, a.steps // 4))
            e = sg.evaluate(P, band=b, ref=ref)
            rows.append({"band": b, "ms": bms, "queries_per_s": K / (bms * 1e-3), "frac_exact": e["frac_exact"],
                         "max_err": e["max_err"], "mean_err": e["mean_err"], "missed_band": e["missed_band"],
                         "max_err_over_diag": e["max_err"] / float((V.max(0)[0] - V.min(0)[0]).norm())})
    return {
        "band_sweep": rows,
        "queries_per_s": K / (ms * 1e-3), "ms_per_step": ms, "band": band,
        "train": {"samples": 1 << 24, "steps": hist["steps"], "batch": hist["batch"], "train_s": hist["train_s"],
                  "losses": hist["losses"]},
        "errors_vs_bvh": ev, "errors_vs_fp64": fp64, "bbox_diag": diag,
        "meets_max_err_1e-5_diag": ev["max_err"] <= 1e-5 * diag,
        "roofline": sg_roofline(infer_ms, tflops, K),
    }


def sg_roofline(infer_ms, tflops, K):
    """Roofline of the surrogate's inference kernel: bound by the hash
    table's gathers beyond L2 (measured FETCH_SIZE + WRITE_SIZE of
    k_sg_infer in its committed counter summary, over the HIP-event-timed
    launch), with the MLP's MFMA rate beside it (measured busy share and
    BF16 FLOPs from the same summary; the FLOP count of the network per
    query when the summary is absent)."""
    f = profile_figures("k_sg_infer_sphere_union_counters")
    traffic = f.get("infer_hbm_bytes_per_launch")
    r = {"bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": None, "traffic": traffic,
         "kernel": "k_sg_infer", "kernel_ms": infer_ms,
         "profile": "profiles/k_sg_infer_sphere_union_counters.md" if f else None,
         "gather_bytes_per_query": 12 + 512 + 4}
    if traffic:
        r["achieved"] = traffic / (infer_ms * 1e-3) / 1e9
        r["frac"] = r["achieved"] / HBM_PEAK_GBS
    mf = f.get("infer_mfma_flops_per_launch")
    achieved = mf / (infer_ms * 1e-3) / 1e12 if mf else tflops
    r["mfma"] = {"achieved": achieved, "peak": 2500.0, "unit": "TFLOP/s", "frac": achieved / 2500.0,
                 "busy_percent": f.get("infer_mfma_busy_percent"),
                 "source": "SQ_INSTS_VALU_MFMA_MOPS_BF16 x 512" if mf else "network FLOPs per query x queries"}
    for k in ("infer_l2_hit_rate", "infer_kernel_ms_profiled"):
        if k in f:
            r[k[6:]] = f[k]
    return r


def bench_surrogate(a):
    """The surrogate + fallback router as its own line (value = routed
    queries/s), next to the plain BVH on the same queries."""
    import torch

    from atom3d_amd import MeshBVH, datasets

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    V, F = datasets.sphere_union()
    mesh = MeshBVH(V.to(dev), F.to(dev))
    P = datasets.uniform_points(a.queries, seed=7).to(dev)
    r = measure_surrogate(mesh, V, P, a, fp64_check=not a.no_cpu, sweep=True)
    for _ in range(a.warmup):
        mesh._bvh.udf(P)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        mesh._bvh.udf(P)
    torch.cuda.synchronize()
    plain_ms = (time.perf_counter() - t0) * 1e3 / a.steps
    res = {
        "metric": "surrogate+fallback UDF queries/s at 4M tris (band 0.01)",
        "value": r["queries_per_s"], "unit": "queries/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": r["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16", "data": "synthetic",
        "config": {"workload": f"hash grid (16 x 2^19 x 2 fp16) + MLP 32-64-64-1 (bf16 MFMA) trained on 2^24 "
                               f"BVH-labelled samples of the union of 200 icospheres (4,096,000 faces), "
                               f"{a.queries} fresh uniform queries in [-1,1]^3, band 0.01",
                   "parallelism": "single GPU"},
        "plain_bvh_queries_per_s": a.queries / (plain_ms * 1e-3),
        "speedup_vs_plain_bvh": plain_ms / r["ms_per_step"],
        **{k: r[k] for k in ("train", "errors_vs_bvh", "errors_vs_fp64", "bbox_diag", "meets_max_err_1e-5_diag",
                             "roofline")},
        "band_sweep": [dict(row, speedup_vs_plain_bvh=plain_ms / row["ms"]) for row in r["band_sweep"]],
        "cpu_baseline": None,
    }
    print(json.dumps(res), flush=True)

def bench_batch(a):
    """256 meshes of log-uniform size (1k-200k faces) with log-uniform query
    sets (2^14-2^20) in their padded boxes, seed 3: one step = one graph
    replay (mesh-major Morton ordering of all queries + one persistent walk
    launch with the autotuned shape)."""
    import torch

    from atom3d_amd import datasets
    from atom3d_amd.batch import BatchedUDF

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    meshes = datasets.uneven_batch(256, seed=3)
    tb = time.perf_counter()
    bu = BatchedUDF([(V, F) for V, F, _ in meshes], device=dev)
    torch.cuda.synchronize()
    build_s = time.perf_counter() - tb
    P, counts = datasets.batch_queries(meshes, seed=3, device=dev)
    K = int(P.shape[0])
    plan = bu.plan(P, counts)
    ta = time.perf_counter()
    from atom3d_amd.batch import autotune_cache_path, load_tuned

    tune_cached = load_tuned(autotune_cache_path(), bu.tune_key(plan)) is not None
    best, table = bu.autotune(plan, reps=2)
    tune_s = time.perf_counter() - ta
    g = bu.capture(plan, best)
    for _ in range(a.warmup):
        g.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        g.replay()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / a.steps
    walk_ms = min(t for _, t in table)
    # per-mesh launches on the same GPU, for comparison
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    o = 0
    for b, c in zip(bu.bvhs, counts):
        b.udf(P[o:o + c])
        o += c
    torch.cuda.synchronize()
    per_mesh_ms = (time.perf_counter() - t1) * 1e3
    cpu = None
    n_node = n_tri = None
    if not a.no_cpu:
        import oracle

        cores = len(os.sched_getaffinity(0))
        rng = np.random.default_rng(0)
        pick = sorted(rng.choice(len(meshes), 16, replace=False).tolist())
        nq, secs, wn, wt = 0, 0.0, 0.0, 0.0
        off = np.concatenate([[0], np.cumsum(counts)])
        for m in pick:
            V, F, _ = meshes[m]
            sah = oracle.SahBvh(V.numpy(), F.numpy())
            sub = P[off[m]:off[m] + min(counts[m], 1 << 13)].cpu().numpy()
            tq = time.perf_counter()
            _, _, _, _, nv, nt = sah.udf(sub, threads=cores, counts=True)
            secs += time.perf_counter() - tq
            nq += len(sub)
            wn += float(nv.sum())
            wt += float(nt.sum())
        n_node, n_tri = wn / nq, wt / nq
        cpu = {"value": nq / secs, "unit": "queries/s", "cores": cores, "kind": "port",
               "sample": f"16 of the 256 meshes (seeded choice), up to 8192 of each one's queries ({nq} in all); "
                         f"the reference SAH walk restated in C++ (OpenMP, {cores} threads), meshes one after "
                         f"another; SAH builds not included"}
    if n_node is None:
        n_node, n_tri = 60.0, 50.0
    bq = 44 + 32 * n_node + 36 * n_tri
    res = {
        "metric": "batched uneven UDF queries/s (256 meshes, one graph launch)",
        "value": K / (ms * 1e-3), "unit": "queries/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": f"256 meshes (noisy icospheres / sphere unions, faces log-uniform 1k-200k, seed 3; "
                               f"{sum(bu.n_faces)} faces), query sets log-uniform 2^14-2^20 in each padded box "
                               f"({K} queries)", "faces": sum(bu.n_faces), "queries": K,
                   "parallelism": "single GPU", "autotuned": best, "autotune_s": tune_s,
                   "autotune_from_cache": tune_cached, "bvh_build_s": build_s},
        "per_mesh_launches_queries_per_s": K / (per_mesh_ms * 1e-3),
        "autotune_table": [[c, t] for c, t in table],
        "roofline": roofline("k_udf_batch", walk_ms, "k_udf_batch_uneven256_counters",
                             (bq * K, f"reference algorithm: B_q = 44 + 32 N_node + 36 N_tri = {bq:.0f} B per "
                                      f"query (oracle counts on a sample of the meshes)")),
        "cpu_baseline": cpu,
    }
    g.close()
    print(json.dumps(res), flush=True)


def bench_ray(a):
    """Nearest-hit rays on the 4.1M-face union of icospheres: 2^24 rays,
    origins uniform in [-1,1]^3, directions uniform on the sphere (seed 1);
    one step = a3d_ray_nearest over all of them.  Also timed: the parity SDF
    (8 directions of all-hit passes + crossing counts) of 2^20 points."""
    import torch

    from atom3d_amd import MeshBVH, datasets

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    V, F = datasets.sphere_union()
    mesh = MeshBVH(V.to(dev), F.to(dev))
    K = a.queries
    rng = np.random.default_rng([1, 2])
    O = torch.from_numpy(rng.uniform(-1, 1, (K, 3)).astype(np.float32)).to(dev)
    Dn = rng.normal(size=(K, 3))
    D = torch.from_numpy((Dn / np.linalg.norm(Dn, axis=1, keepdims=True)).astype(np.float32)).to(dev)
    del Dn
    bvh = mesh._bvh
    ms, (hit, t, f, _, _) = _timed(lambda: bvh.ray_nearest(O, D), a)
    P = datasets.uniform_points(1 << 20, seed=2).to(dev)
    steps, warm = a.steps, a.warmup
    a.steps, a.warmup = max(2, a.steps // 5), 1
    ms_par, _ = _timed(lambda: mesh.sdf(P, mode="parity"), a)
    a.steps, a.warmup = steps, warm
    cpu = None
    n_node, n_tri = 123.0, 10.0
    if not a.no_cpu:
        import oracle

        cores = len(os.sched_getaffinity(0))
        tb = time.perf_counter()
        sah = oracle.SahBvh(V.numpy(), F.numpy())
        sah_s = time.perf_counter() - tb
        n = min(a.cpu_sample, K)
        Oc, Dc = O[:n].cpu().numpy(), D[:n].cpu().numpy()
        tq = time.perf_counter()
        rh, rt, rf, (wn, wt) = sah.ray(Oc, Dc, threads=cores, work=True)
        cpu_s = time.perf_counter() - tq
        n_node, n_tri = wn / n, wt / n
        agree = float(np.mean(rf == f[:n].cpu().numpy()))
        cpu = {"value": n / cpu_s, "unit": "rays/s", "cores": cores, "kind": "port",
               "sample": f"the first {n} of the {K} rays; the reference's nearest-hit walk restated in C++ "
                         f"(OpenMP, {cores} threads) on the reference SAH tree (build {sah_s:.1f} s, not "
                         f"included); same face for {agree:.6f} of them"}
    unit_bytes = 24 + 36 * n_node + 40 * n_tri
    out = {
        "metric": "nearest-hit rays/s at 4M tris", "value": K / (ms * 1e-3), "unit": "rays/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": f"MeshBVH.intersect_ray (nearest) on the union of 200 icospheres (4,096,000 faces, "
                               f"seed 1): {K} rays, origins uniform in [-1,1]^3, uniform directions",
                   "rays": K, "hit_fraction": float(hit.float().mean()), "parallelism": "single GPU",
                   "parity_sdf_points_per_s": (1 << 20) / (ms_par * 1e-3), "parity_sdf_ms": ms_par,
                   "n_node": n_node, "n_tri": n_tri},
        "roofline": roofline("k_ray_nearest", ms, "k_ray_nearest_sphere_union_counters",
                             (unit_bytes * K, f"reference algorithm per ray: 24 + 36 N_node + 40 N_tri = "
                                              f"{unit_bytes:.0f} B (oracle counts)")),
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def _timed(fn, a):
    import torch

    for _ in range(a.warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        r = fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / a.steps, r


def bench_clip(a):
    """The reference demo's workload (demo.ipynb cells 1-5): unit icosphere
    with 1 subdivision, octree over the mesh bounds to level 9 (512^3), then
    intersect_aabb on every candidate voxel: mode 1 (SAT pairs) and mode 3
    (pairs clipped to polygons with centroid and area).  One step = mode 3."""
    import torch

    from atom3d_amd import MeshBVH, OctreeIndexer, datasets

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    v, f = datasets.icosphere(1, 1.0)
    V, F = torch.from_numpy(v.astype(np.float32)), torch.from_numpy(f.astype(np.int32))
    m = MeshBVH(V.to(dev), F.to(dev))
    oc = OctreeIndexer(max_level=9, bounds=m.get_bounds(), device=dev)
    cand = oc.octree_traverse(m)
    lo, hi = oc.cube_aabb_level(cand)
    lo, hi = lo.contiguous(), hi.contiguous()
    ms1, r1 = _timed(lambda: m.intersect_aabb(lo, hi, mode=1), a)
    ms3, r3 = _timed(lambda: m.intersect_aabb(lo, hi, mode=3), a)
    # Voxelizer(bvh, CubeGrid(512, mesh bounds)).voxelize_with_faces(): every
    # cell of the faces' block (here the whole 512^3 grid) SAT-tested
    from atom3d_amd import CubeGrid
    from atom3d_amd.apps import Voxelizer

    vz = Voxelizer(m, CubeGrid(512, bounds=m.get_bounds(), device=dev))
    msv, vf = _timed(vz.voxelize_with_faces, a)
    vox_leg = {"ms": msv, "cells_tested": 512 ** 3, "cells_per_s": 512 ** 3 / (msv * 1e-3),
               "voxels": int(vf.voxel_coords.shape[0]), "pairs": int(vf.face_indices.shape[0]),
               "octree_voxels": int(cand.shape[0])}
    n_pairs = int(r1.aabb_ids.shape[0])
    n_poly = int(r3.poly_counts.shape[0])
    cpu = None
    if not a.no_cpu:
        import oracle

        cores = len(os.sched_getaffinity(0))
        sah = oracle.SahBvh(v.astype(np.float32), f.astype(np.int32))
        k = min(len(cand), 200000)
        bl, bh = lo[:k].cpu().numpy(), hi[:k].cpu().numpy()
        tq = time.perf_counter()
        _, pa, pf = sah.aabb(bl, bh)
        tris9 = v.astype(np.float32)[f.astype(np.int64)].reshape(-1, 9)
        out = oracle.sat_clip(bl, bh, tris9, pa.astype(np.int64), pf.astype(np.int64), mode=3)
        cpu_s = time.perf_counter() - tq
        cpu = {"value": len(pa) / cpu_s, "unit": "clipped pairs/s", "cores": 1, "kind": "port",
               "sample": f"the first {k} of the {len(cand)} candidate voxels ({len(pa)} SAT pairs); the reference "
                         f"SAH pair walk and Sutherland-Hodgman clip restated in C++ (single thread)"}
        _ = out
    per_pair = 24 + 36 + 16 + 1 + 4 + 96 + 12 + 4
    res = {
        "metric": "intersect_aabb mode 3 clipped pairs/s (demo icosphere, 512^3)", "value": n_pairs / (ms3 * 1e-3),
        "unit": "pairs/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": ms3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": f"icosphere(1) (80 faces), octree level 9 over its bounds: {len(cand)} candidate "
                               f"voxels, {n_pairs} SAT pairs, {n_poly} non-empty polygons (demo known answer "
                               f"1,172,044)", "parallelism": "single GPU"},
        "mode1_ms": ms1, "mode1_pairs_per_s": n_pairs / (ms1 * 1e-3), "voxelize_with_faces": vox_leg,
        "roofline": roofline("whole mode-3 call", ms3, "clip_demo512_counters",
                             (per_pair * n_pairs, f"{per_pair} B per clipped pair (box, triangle, outputs)")),
        "cpu_baseline": cpu,
    }
    print(json.dumps(res), flush=True)


def bench_c1(a):
    """The reference's CPU-runnable case: noisy icosphere (5 subdivisions,
    r 0.5, N(0, 0.01) radial noise, seed 0; 20,480 faces), 131,072 uniform
    queries in [-1,1]^3.  One step = UDF (Morton order + walk); SDF timed too."""
    import torch

    from atom3d_amd import MeshBVH, datasets

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    V, F = datasets.noisy_icosphere()
    m = MeshBVH(V.to(dev), F.to(dev))
    P = datasets.uniform_points(131072, seed=0).to(dev)
    K = P.shape[0]
    a.steps = max(a.steps, 50)
    cold_ms, _ = _timed(lambda: m._bvh.udf(P), a)  # a mesh's first calls: the plain per-lane walk
    m._bvh.seed_grid()  # steady state, like the headline: the distance grid built once per mesh
    ms, _ = _timed(lambda: m._bvh.udf(P), a)
    ms_sdf, _ = _timed(lambda: m.sdf(P), a)
    cpu = None
    n_node = n_tri = None
    if not a.no_cpu:
        import oracle

        cores = len(os.sched_getaffinity(0))
        sah = oracle.SahBvh(V.numpy(), F.numpy())
        Pn = P.cpu().numpy()
        t_all = median_of(lambda: sah.udf(Pn, threads=cores))
        t_one = median_of(lambda: sah.udf(Pn[:16384], threads=1))
        _, _, _, _, nv, nt = sah.udf(Pn, threads=cores, counts=True)
        n_node, n_tri = float(nv.mean()), float(nt.mean())
        cpu = {"value": K / t_all, "unit": "queries/s", "cores": cores, "kind": "port",
               "single_thread_value": 16384 / t_one, "cpu_model": cpu_model(),
               "protocol": "one warm-up run, then the median of 5",
               "sample": f"all cores: all {K} queries; single thread: the first 16384; the reference SAH build "
                         f"and walk restated in C++ (OpenMP)"}
    if n_node is None:
        n_node, n_tri = 40.0, 30.0
    bq = 44 + 32 * n_node + 36 * n_tri
    res = {
        "metric": "closest-point queries/s, noisy icosphere (20,480 tris)", "value": K / (ms * 1e-3),
        "unit": "queries/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": "noisy icosphere (sub 5, r 0.5, sigma 0.01, seed 0; 20,480 faces), 131,072 uniform "
                               "queries in [-1,1]^3 (seed 0); per-lane walk bounded by the mesh's distance grid "
                               "(built once per mesh, like the tree)", "parallelism": "single GPU"},
        "sdf_queries_per_s": K / (ms_sdf * 1e-3),
        "without_grid_queries_per_s": K / (cold_ms * 1e-3),
        "gpu_over_cpu": (K / (ms * 1e-3)) / cpu["value"] if cpu else None,
        "roofline": roofline("whole step (launch-bound size)", ms, "udf_icosphere_c1_counters",
                             (bq * K, f"reference algorithm: B_q = 44 + 32 N_node + 36 N_tri = {bq:.0f} B per "
                                      f"query (oracle counts)")),
        "cpu_baseline": cpu,
    }
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
