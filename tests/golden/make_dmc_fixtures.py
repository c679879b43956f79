"""Writes tests/golden/dmc_fixtures.npz: input/output vectors of the
reference's SparseDiffDMC (Atom3D atom3d/mesh_extractor/sparse_diffdmc.py,
weighted_avg mode) on small CPU inputs, forward outputs plus the gradients
of a fixed random linear loss.

The reference is pure PyTorch here; only its three source files
(constants.py, mesh_extractor/tables.py, mesh_extractor/sparse_diffdmc.py)
are loaded, under a private package name, so the package's __init__ (which
JIT-builds CUDA extensions) is never executed.  Run in a container that has
the reference at /root/reference:  python tests/golden/make_dmc_fixtures.py
"""
import importlib.util
import os
import sys
import types

import numpy as np
import torch

REF = "/root/reference/atom3d"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "dmc_fixtures.npz")
HERE = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))


def _load_reference():
    for name, path in (("_a3dref", REF), ("_a3dref.mesh_extractor", REF + "/mesh_extractor")):
        m = types.ModuleType(name)
        m.__path__ = [path]
        sys.modules[name] = m
    for name, f in (("_a3dref.constants", "constants.py"), ("_a3dref.mesh_extractor.tables", "mesh_extractor/tables.py"),
                    ("_a3dref.mesh_extractor.sparse_diffdmc", "mesh_extractor/sparse_diffdmc.py")):
        spec = importlib.util.spec_from_file_location(name, os.path.join(REF, f))
        mod = importlib.util.module_from_spec(spec)
        sys.modules[name] = mod
        spec.loader.exec_module(mod)
    return sys.modules["_a3dref.mesh_extractor.sparse_diffdmc"].SparseDiffDMC


def _cases():
    sys.path.insert(0, HERE)
    from atom3d_amd.datasets import shapes_sdf  # noqa: E402
    from oracle.dmc import grid_band  # noqa: E402

    cases = []
    # 2 spheres + 1 torus, 16^3 band |sdf| < 2h
    cubes, s, ci = grid_band(shapes_sdf, 16, 2 * 2.0 / 16)
    cases.append(("shapes16", cubes, s, ci, 16))
    # random field on a dense 6^3 grid: many ambiguous faces and check flips
    r = 6
    g = np.arange(r)
    c = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    cid = np.stack([((c[:, 0] + (i & 1)) * (r + 1) + c[:, 1] + ((i >> 1) & 1)) * (r + 1) + c[:, 2] + ((i >> 2) & 1)
                    for i in range(8)], 1)
    rng = np.random.default_rng(7)
    field = rng.standard_normal((r + 1) ** 3)
    cases.append(("random6", c.astype(np.int64), field, cid.astype(np.int64), r))
    # sparse subset (70 %) of the same cubes with re-indexed corners: edges
    # shared by fewer than 4 cubes, missing check neighbours
    keep = rng.random(len(c)) < 0.7
    used, inv = np.unique(cid[keep].reshape(-1), return_inverse=True)
    cases.append(("random6_sparse", c[keep].astype(np.int64), field[used], inv.reshape(-1, 8).astype(np.int64), r))
    return cases


def main():
    SparseDiffDMC = _load_reference()
    out = {}
    for name, cubes, s, ci, res in _cases():
        N, M = len(cubes), len(s)
        rng = np.random.default_rng(len(name) * 100000 + N)
        alpha = rng.standard_normal((N, 8))
        beta = rng.standard_normal((N, 12))
        deform = rng.uniform(-0.1, 0.1, (M, 3)) * (2.0 / res)
        feat = rng.standard_normal((M, 3))
        out.update({f"{name}/voxel_coords": cubes.astype(np.int32), f"{name}/cube_idx": ci.astype(np.int32),
                    f"{name}/res": np.array(res), f"{name}/sdf": s, f"{name}/alpha": alpha, f"{name}/beta": beta,
                    f"{name}/deform": deform, f"{name}/v_features": feat})
        for dt, tag in ((torch.float64, "f64"), (torch.float32, "f32")):
            for variant in ("plain", "full"):
                key = f"{name}_{tag}_{variant}"
                m = SparseDiffDMC(device="cpu", dtype=dt)
                sd = torch.tensor(s, dtype=dt, requires_grad=True)
                kw = {}
                leaves = {"sdf": sd}
                if variant == "full":
                    for k, v in (("alpha", alpha), ("beta", beta), ("deform", deform), ("v_features", feat)):
                        leaves[k] = kw[k] = torch.tensor(v, dtype=dt, requires_grad=True)
                o = m(torch.tensor(cubes), sd, torch.tensor(ci), res, isovalue=0.0, **kw)
                loss = 0
                outs = {"vertices": o.vertices, "L_dev": o.L_dev, "p_alpha": o.p_alpha, "n_alpha": o.n_alpha}
                if o.v_features is not None:
                    outs["v_features"] = o.v_features
                for k, t in outs.items():
                    w = torch.tensor(rng.standard_normal(tuple(t.shape)), dtype=dt)
                    out[f"{key}/W_{k}"] = w.numpy()
                    loss = loss + (w * t).sum()
                loss.backward()
                for k, t in outs.items():
                    out[f"{key}/out_{k}"] = t.detach().numpy()
                out[f"{key}/out_faces"] = o.faces.numpy()
                for k, t in leaves.items():
                    if t.grad is not None:
                        out[f"{key}/grad_{k}"] = t.grad.numpy()
                if variant == "plain" and tag == "f64":
                    m.eval()
                    oe = m(torch.tensor(cubes), sd.detach(), torch.tensor(ci), res, inference_mode="weighted_avg")
                    out[f"{key}/eval_faces"] = oe.faces.numpy()
                    out[f"{key}/eval_vertices"] = oe.vertices.numpy()
                print(key, "V", tuple(o.vertices.shape), "F", tuple(o.faces.shape), "E", tuple(o.L_dev.shape),
                      flush=True)
    np.savez_compressed(OUT, **out)
    print(OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
