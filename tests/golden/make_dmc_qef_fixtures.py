"""Writes tests/golden/dmc_qef_fixtures.npz: input/output vectors of the
reference's SparseDiffDMC in QEF mode (Atom3D
atom3d/mesh_extractor/sparse_diffdmc.py, eval mode, inference_mode="qef":
corner gradients :450-513, crossing normals :617-655, QEF :289-389, eval
split :802-810) on the same small CPU inputs as dmc_fixtures.npz.

Variants per input and dtype: "plain" (no deform, gradients estimated),
"deform" (random corner deformation and features, gradients estimated) and
"given" (caller-supplied corner gradients, as from a BVH: a random field
with some zero rows, which take the edge-direction fallback).

Loads the reference's three pure-PyTorch source files the same way as
make_dmc_fixtures.py.  Run in a container that has the reference at
/root/reference:  python tests/golden/make_dmc_qef_fixtures.py
"""
import os
import sys

import numpy as np
import torch

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from make_dmc_fixtures import _cases, _load_reference  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "dmc_qef_fixtures.npz")


def main():
    SparseDiffDMC = _load_reference()
    out = {}
    for name, cubes, s, ci, res in _cases():
        N, M = len(cubes), len(s)
        rng = np.random.default_rng(len(name) * 7919 + M)
        deform = rng.uniform(-0.15, 0.15, (M, 3)) * (2.0 / res)
        feat = rng.standard_normal((M, 2))
        cg = rng.standard_normal((M, 3))
        cg[rng.random(M) < 0.2] = 0.0
        out.update({f"{name}/voxel_coords": cubes.astype(np.int32), f"{name}/cube_idx": ci.astype(np.int32),
                    f"{name}/res": np.array(res), f"{name}/sdf": s, f"{name}/deform": deform,
                    f"{name}/v_features": feat, f"{name}/corner_gradients": cg})
        for dt, tag in ((torch.float64, "f64"), (torch.float32, "f32")):
            for variant in ("plain", "deform", "given"):
                key = f"{name}_{tag}_{variant}"
                m = SparseDiffDMC(device="cpu", dtype=dt)
                m.eval()
                kw = {}
                if variant == "deform":
                    kw = dict(deform=torch.tensor(deform, dtype=dt), v_features=torch.tensor(feat, dtype=dt))
                if variant == "given":
                    kw = dict(corner_gradients=torch.tensor(cg, dtype=dt))
                with torch.no_grad():
                    o = m(torch.tensor(cubes), torch.tensor(s, dtype=dt), torch.tensor(ci), res, **kw)
                for k in ("vertices", "faces", "L_dev", "p_alpha", "n_alpha", "v_features"):
                    t = getattr(o, k)
                    if t is not None:
                        out[f"{key}/out_{k}"] = t.numpy()
                print(key, "V", tuple(o.vertices.shape), "F", tuple(o.faces.shape), flush=True)
    np.savez_compressed(OUT, **out)
    print(OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
