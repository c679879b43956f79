"""BVHAccelerator: device BVH build and queries over the HIP library.

Mirrors Atom3D atom3d/kernels/bvh.py:54-207 (same constructor arguments, same
method names and return tuples; ray_intersect / ray_all_hit :119-181).  Differences, all deliberate:

* the tree is built on the GPU (Morton LBVH) instead of on the host, and is
  held as one opaque byte blob (``self.blob``) instead of ``nodes``/``triangles``
  tensors;
* the closest face is tree independent (minimum FP32 squared distance, ties to
  the lowest face index) and ``uvw`` is zero where no face was found (the
  reference leaves it uninitialised);
* the box queries (``aabb_hit``, ``aabb_intersect``, ``block_cells``,
  ``cells_sat_pairs`` and the octree level test) run on a second tree,
  ``box_blob``: the reference's own binned-SAH tree rebuilt on the device
  (same leaves), because the reference keeps a face when its LEAF box
  overlaps the query box and the SAT passes, which depends on the leaves;
  it is built on the first box query;
* ``aabb_intersect`` never truncates: it counts pairs first, then writes them
  sorted by (box, face) (the reference appends with an atomic counter, capped
  at min(100*K, 5e7) pairs, in nondeterministic order);
* ``n_primitives_per_leaf`` sets the box tree's leaves (reference default 4);
  the distance/ray tree uses 8 unless it is given: with tree-independent
  results it only trades node visits against leaf tests, and 8 is faster here;
* ``treelet_passes`` (default 0 = plain LBVH, asynchronous build) sets how
  many SAH treelet-restructuring passes refine the device tree.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib


class BVHAccelerator:
    TREELET_PASSES = 0

    def __init__(self, vertices: torch.Tensor, faces: torch.Tensor, n_primitives_per_leaf: int | None = None,
                 treelet_passes: int | None = None):
        _lib.require_cuda(vertices, "vertices")
        self.device = vertices.device
        self.vertices = vertices.contiguous().float()
        self.faces = faces.to(self.device).contiguous().int()
        self.num_faces = int(self.faces.shape[0])
        self.leaf_max = 8 if n_primitives_per_leaf is None else int(n_primitives_per_leaf)
        self.box_leaf = 4 if n_primitives_per_leaf is None else int(n_primitives_per_leaf)
        self._box_blob = None
        self.box_depth = None
        self.treelet_passes = self.TREELET_PASSES if treelet_passes is None else int(treelet_passes)
        if self.num_faces > 0:
            fmin, fmax = int(self.faces.min()), int(self.faces.max())
            if fmin < 0 or fmax >= self.vertices.shape[0]:
                raise ValueError(
                    f"faces index out of range [0, {self.vertices.shape[0]}): min {fmin}, max {fmax}"
                )
        blob_b, ws_b = ctypes.c_int64(), ctypes.c_int64()
        _lib.call("a3d_bvh_sizes", self.num_faces, self.leaf_max, ctypes.byref(blob_b), ctypes.byref(ws_b))
        with torch.cuda.device(self.device):
            self.blob = torch.empty(blob_b.value, dtype=torch.uint8, device=self.device)
            ws = torch.empty(ws_b.value, dtype=torch.uint8, device=self.device)
            _lib.call(
                "a3d_bvh_build_ex",
                _lib.ptr(self.vertices), self.vertices.shape[0], _lib.ptr(self.faces), self.num_faces,
                self.leaf_max, self.treelet_passes, _lib.ptr(self.blob), blob_b.value, _lib.ptr(ws), ws_b.value,
                _lib.stream(self.device),
            )
            del ws

    @property
    def box_blob(self) -> torch.Tensor:
        """The box queries' tree: the reference's binned-SAH tree with
        ``box_leaf`` triangles per leaf (kernels/bvh_kernels.cu:616-925),
        built on the device on first use."""
        if self._box_blob is None:
            blob_b, ws_b = ctypes.c_int64(), ctypes.c_int64()
            _lib.call("a3d_sah_sizes", self.num_faces, self.box_leaf, ctypes.byref(blob_b), ctypes.byref(ws_b))
            depth = ctypes.c_int32()
            with torch.cuda.device(self.device):
                blob = torch.empty(blob_b.value, dtype=torch.uint8, device=self.device)
                ws = torch.empty(ws_b.value, dtype=torch.uint8, device=self.device)
                _lib.call("a3d_sah_build", _lib.ptr(self.vertices), self.vertices.shape[0], _lib.ptr(self.faces),
                          self.num_faces, self.box_leaf, _lib.ptr(blob), blob_b.value, _lib.ptr(ws), ws_b.value,
                          ctypes.byref(depth), _lib.stream(self.device))
                del ws
            self._box_blob = blob
            self.box_depth = int(depth.value)
        return self._box_blob

    def box_tree(self):
        """The box tree on the host, for comparisons with the reference's:
        (order int32 [F] — original face ids in the tree's triangle order,
        leaves int64 [L, 2] — (start, end) triangle ranges, leaf boxes
        float32 [L, 6] (min, max), internal-node boxes float32 [N, 6] (the
        root's is the union of its children's: records hold child boxes
        only), depth)."""
        import numpy as np

        raw = self.box_blob.cpu().numpy()
        hdr = raw[:256].view(np.int32)
        n_tris, n_nodes, cap = int(hdr[1]), int(hdr[2]), int(hdr[3])
        nodes_off, tris_off = int(raw[16:24].view(np.int64)[0]), int(raw[24:32].view(np.int64)[0])
        rec = raw[nodes_off:nodes_off + 64 * cap].view(np.float32).reshape(cap, 16)
        refs = rec[:, 12:14].view(np.int32)
        order = raw[tris_off:tris_off + 48 * n_tris].view(np.float32).reshape(n_tris, 12)[:, 9].view(np.int32).copy()
        depth = int(hdr[17])
        if n_tris == 0:
            return order, np.zeros((0, 2), np.int64), np.zeros((0, 6), np.float32), np.zeros((0, 6), np.float32), depth
        # records reachable from the root (reserved records of small subtrees may be unused)
        seen = np.zeros(cap, bool)
        frontier = np.array([0])
        seen[0] = True
        while len(frontier):
            kids = refs[frontier].reshape(-1)
            kids = kids[kids >= 0]
            seen[kids] = True
            frontier = kids
        live = np.nonzero(seen)[0]
        boxes = rec[live, :12].reshape(-1, 2, 6)
        r = refs[live]
        leaf = (r < 0) & (boxes[:, :, 0] <= boxes[:, :, 3])
        enc = ~r[leaf].astype(np.int64)
        leaves = np.stack([enc >> 3, (enc >> 3) + (enc & 7) + 1], 1)
        root = np.concatenate([np.minimum(boxes[0, 0, :3], boxes[0, 1, :3]),
                               np.maximum(boxes[0, 0, 3:], boxes[0, 1, 3:])]) if n_tris > self.box_leaf else None
        inner = boxes[r >= 0]
        if root is not None:
            inner = np.concatenate([root[None], inner])
        return order, leaves, boxes[leaf], inner, depth

    def info(self) -> dict:
        out = (ctypes.c_int64 * 4)()
        with torch.cuda.device(self.device):
            _lib.call("a3d_bvh_info", _lib.ptr(self.blob), out, _lib.stream(self.device))
        return {"n_tris": out[0], "n_nodes": out[1], "depth": out[2], "leaf_max": out[3]}

    def check_tree(self) -> dict:
        """Copy the blob to the host and verify its structure: every internal
        reference points at a written node, leaves partition the triangles,
        each child box contains its triangles.  Raises AssertionError."""
        import numpy as np

        info = self.info()
        raw = self.blob.cpu().numpy()
        hdr = raw[:256].view(np.int32)
        n_tris, n_nodes, cap = int(hdr[1]), int(hdr[2]), int(hdr[3])
        assert hdr[0] == 0x41334442 and n_tris == self.num_faces, "bad header"
        nodes_off, tris_off = raw[16:24].view(np.int64)[0], raw[24:32].view(np.int64)[0]
        nodes = raw[nodes_off:nodes_off + 64 * cap].view(np.float32).reshape(cap, 16)[:n_nodes]
        refs = nodes[:, 12:14].view(np.int32)
        tris = raw[tris_off:tris_off + 48 * n_tris].view(np.float32).reshape(n_tris, 12)
        ids = tris[:, 9].view(np.int32)
        assert 1 <= n_nodes <= cap
        if n_tris:
            assert np.array_equal(np.sort(ids), np.arange(n_tris)), "triangle ids are not a permutation"
        covered = np.zeros(max(n_tris, 1), np.int32)
        seen = np.zeros(n_nodes, np.int32)
        seen[0] = 1
        tv = tris[:, :9].reshape(-1, 3, 3) if n_tris else None
        for side in range(2):
            r = refs[:, side]
            box = nodes[:, side * 6:side * 6 + 6]
            internal = r >= 0
            assert np.all(r[internal] < n_nodes) and np.all(r[internal] > 0), "internal ref out of range"
            np.add.at(seen, r[internal], 1)
            if n_tris == 0:
                continue
            leaf = (~internal) & (box[:, 0] <= box[:, 3])  # empty children carry inverted boxes
            enc = ~r[leaf]
            start, cnt = enc >> 3, (enc & 7) + 1
            assert np.all(start + cnt <= n_tris), "leaf range out of range"
            tri_idx = np.repeat(start, cnt) + (np.arange(cnt.sum()) - np.repeat(np.cumsum(cnt) - cnt, cnt))
            np.add.at(covered, tri_idx, 1)
            owner = np.repeat(np.nonzero(leaf)[0], cnt)
            t = tv[tri_idx]
            assert np.all(t.min(1) >= box[owner, :3]) and np.all(t.max(1) <= box[owner, 3:]), "leaf box"
        assert np.all(seen == 1), "every node is referenced exactly once (root by the walk)"
        if n_tris:
            assert np.all(covered[:n_tris] == 1), "leaves do not partition the triangles"
            self._check_obbs(raw, refs, tv, n_nodes)
        return info

    @staticmethod
    def _check_obbs(raw, refs, tv, n_nodes):
        """Every child's oriented box holds its subtree's vertices on each of
        its three axes (n, t1, t2 = n x t1 in float32): lo <= a . v <= hi in
        FP64, |a| <= 1, axes orthogonal to float precision; children above
        256 triangles have none."""
        import numpy as np

        off = int(raw[80:88].view(np.int64)[0])
        assert off > 0, "blob has no oriented boxes"
        ob = raw[off:off + 96 * n_nodes].view(np.float32).reshape(n_nodes, 2, 12)
        memo = {}

        def tris_of(ref):
            if ref < 0:
                enc = ~ref
                return np.arange(enc >> 3, (enc >> 3) + (enc & 7) + 1)
            if ref not in memo:
                memo[ref] = np.concatenate([tris_of(int(refs[ref, 0])), tris_of(int(refs[ref, 1]))])
            return memo[ref]

        for node in range(n_nodes - 1, -1, -1):  # children have larger indices: memo fills bottom-up
            for side in range(2):
                c = ob[node, side]
                n, t1 = c[0:3], c[4:7]
                if not np.any(n):
                    continue
                t2 = np.array([n[1] * t1[2] - n[2] * t1[1], n[2] * t1[0] - n[0] * t1[2],
                               n[0] * t1[1] - n[1] * t1[0]], np.float32)
                axes = np.stack([n, t1, t2]).astype(np.float64)
                assert (np.linalg.norm(axes, axis=1) <= 1.0).all(), "axis longer than 1"
                g = axes @ axes.T
                assert np.abs(g - np.diag(np.diag(g))).max() < 1e-6, "axes not orthogonal"
                lo = np.array([c[3], c[7], c[10]], np.float64)
                hi = np.array([c[8], c[9], c[11]], np.float64)
                idx = tris_of(int(refs[node, side]))
                assert len(idx) <= 256, "oriented box on a child of more than 256 triangles"
                d = tv[idx].reshape(-1, 3).astype(np.float64) @ axes.T
                assert (d.min(0) >= lo).all() and (d.max(0) <= hi).all(), "oriented box does not hold its triangles"

    def _points(self, points: torch.Tensor) -> torch.Tensor:
        _lib.require_cuda(points, "points")
        if points.device != self.device:
            raise RuntimeError(f"atom3d_amd: points on {points.device}, BVH on {self.device}")
        if points.dim() != 2 or points.shape[1] != 3:
            raise ValueError(f"points must be [K, 3], got {tuple(points.shape)}")
        return points.contiguous().float()

    UDF_LANE = 1
    UDF_DIRECT = 2
    UDF_NO_OBBS = 4
    UDF_SPLIT = 8
    UDF_REDO = 16
    ORDER_MIN_QUERIES = 4096
    SEED_MIN_QUERIES = 1 << 15
    SEED_GRID_RES = 127  # (R+1)^3 = 2^21 grid points
    SEED_AMORTIZE = 4
    SPLIT_MAX_QUERIES = 1 << 15
    LANE_MAX_QUERIES = 1 << 17

    def query_order(self, points: torch.Tensor, by_cost: bool = False) -> torch.Tensor:
        """Morton order of the query points (uint32 permutation stored in an
        int32 tensor).  ``by_cost`` (with the seed grid built): the Morton
        order's 64-query packets regrouped by the grid's per-point walk work,
        heaviest class first (a3d_query_order_work), the order for the
        grid-bounded walk."""
        p = self._points(points)
        K = p.shape[0]
        sg = self.__dict__.get("_seed_grids", {}).get(self.SEED_GRID_RES) if by_cost else None
        work = self.__dict__.get("_seed_work", {}).get(self.SEED_GRID_RES) if sg else None
        name = "a3d_query_order" if work is None else "a3d_query_order_work"
        wsb = ctypes.c_int64()
        _lib.call(name + "_ws_bytes", K, ctypes.byref(wsb))
        with torch.cuda.device(self.device):
            perm = torch.empty(K, dtype=torch.int32, device=self.device)
            ws = torch.empty(max(wsb.value, 1), dtype=torch.uint8, device=self.device)
            if work is None:
                _lib.call(name, _lib.ptr(p), K, _lib.ptr(perm), _lib.ptr(ws), wsb.value, _lib.stream(self.device))
            else:
                _lib.call(name, _lib.ptr(p), K, _lib.ptr(work), self.SEED_GRID_RES, (ctypes.c_float * 3)(*sg[2]),
                          sg[3], _lib.ptr(perm), _lib.ptr(ws), wsb.value, _lib.stream(self.device))
        return perm

    def seed_grid(self, res: int | None = None):
        """(grid, grid_face, lo, h), or None for a mesh with non-finite
        coordinates: exact distances and closest faces at the (R+1)^3 points
        lo + (i, j, k) h (float32 arithmetic, R = ``res``, default
        SEED_GRID_RES; R + 1 a power of two) of a cube holding the mesh with
        a 5 % margin, built once per mesh and resolution: the points are
        generated on the device in a walk order of 4x4x4 bricks along a
        Morton curve (a3d_grid_points), then walked once by the plain wave
        walk, which also records each point's walk work (a3d_bvh_udf_work,
        kept for query_order's cost order).  The source of the grid-bounded
        walk's per-query distance bounds."""
        R = int(res or self.SEED_GRID_RES)
        side = R + 1
        if side < 4 or side & (side - 1):
            raise ValueError(f"seed grid resolution R={R}: R + 1 must be a power of two >= 4")
        grids = self.__dict__.setdefault("_seed_grids", {})
        if R not in grids:
            import math

            import numpy as np

            # the exact box of the referenced vertices, from the blob header (one 256-byte copy)
            hdr = self.blob[:256].cpu().numpy()
            box = hdr[32:56].view(np.float32).astype(np.float64)
            if not np.all(np.isfinite(box)) or not bool(torch.isfinite(self.vertices).all()):
                grids[R] = False  # non-finite coordinates: the plain walk only
                return None
            lo_b, hi_b = box[:3], box[3:]
            ext = float((hi_b - lo_b).max())
            if not math.isfinite(ext):
                grids[R] = False
                return None
            pad = 0.05 * ext + 1e-6
            h = float(np.float32((ext + 2 * pad) / R))
            lo = [float(np.float32(x)) for x in (lo_b - pad).tolist()]
            n = side ** 3
            with torch.cuda.device(self.device):
                st = _lib.stream(self.device)
                pts = torch.empty(n, 3, dtype=torch.float32, device=self.device)
                perm = torch.empty(n, dtype=torch.int32, device=self.device)
                _lib.call("a3d_grid_points", (ctypes.c_float * 3)(*lo), h, side, _lib.ptr(pts), _lib.ptr(perm), st)
                d = torch.empty(n, dtype=torch.float32, device=self.device)
                f = torch.empty(n, dtype=torch.int32, device=self.device)
                c = torch.empty(n, 3, dtype=torch.float32, device=self.device)
                work = torch.empty(n, dtype=torch.int32, device=self.device)
                _lib.call("a3d_bvh_udf_work", _lib.ptr(self.blob), _lib.ptr(pts), n, _lib.ptr(perm), _lib.ptr(d),
                          _lib.ptr(f), _lib.ptr(c), None, 0, _lib.ptr(work), st)
                del pts, perm, c
            self.__dict__.setdefault("_seed_work", {})[R] = work
            grids[R] = (d, f, lo, h)
        return grids[R] or None

    def _grid_pays(self, K: int) -> bool:
        """Whether this call should use the exact-distance grid: built
        already, or (from SEED_MIN_QUERIES queries) once this mesh has
        answered SEED_AMORTIZE grids' worth of queries without it, so that
        one-shot calls never pay for the grid (about a plain walk of
        (R+1)^3 queries) and repeated calls on a mesh get it."""
        grids = self.__dict__.get("_seed_grids", {})
        if self.SEED_GRID_RES in grids:
            return bool(grids[self.SEED_GRID_RES])
        if K < self.SEED_MIN_QUERIES or self.num_faces == 0:
            return False
        seen = self.__dict__.get("_queries_seen", 0)
        self._queries_seen = seen + K
        return seen >= self.SEED_AMORTIZE * (self.SEED_GRID_RES + 1) ** 3

    def udf(self, points: torch.Tensor, flags: int | None = None, order: bool | None = None,
            counters: torch.Tensor | None = None, seed: bool | None = None):
        """Closest point on the mesh: (distances [K], face_ids [K] int32,
        closest_points [K,3], uvw [K,3]).  ``flags`` picks the traversal
        (default: the per-lane walk in input order up to 2^17 queries, else
        the packet walk in Morton order; UDF_LANE, UDF_DIRECT or 0 force one);
        ``order`` (default: K >= 4096) walks the queries in Morton order;
        ``counters`` (int64[5] on the device) accumulates the packet walk's
        work counts; ``seed`` bounds each query's walk from the distance
        grid of ``seed_grid`` and walks the few queries whose bound fell
        short again without one (default, packet or per-lane walk: once the
        grid pays for itself, see ``_grid_pays``; True builds it now).
        Results are identical either way."""
        p = self._points(points)
        K = p.shape[0]
        dist = torch.empty(K, dtype=torch.float32, device=self.device)
        face = torch.empty(K, dtype=torch.int32, device=self.device)
        closest = torch.empty(K, 3, dtype=torch.float32, device=self.device)
        uvw = torch.empty(K, 3, dtype=torch.float32, device=self.device)
        if flags is None:
            # few queries: per-lane walks in input order are latency-bound at
            # a lower floor than the packet walk (whose packets are sparse);
            # the fewest get 4 lanes each (the split walk)
            flags = 0
            if counters is None and K <= self.SPLIT_MAX_QUERIES:
                flags = self.UDF_SPLIT
            elif counters is None and K <= self.LANE_MAX_QUERIES:
                flags = self.UDF_LANE
            if flags and order is None:
                order = False
        if order is None:
            order = self.ORDER_MIN_QUERIES <= K < 2 ** 31
        if seed is None:
            seed = flags in (0, self.UDF_LANE) and counters is None and self._grid_pays(K)
        with torch.cuda.device(self.device):
            sg = self.seed_grid() if seed and K > 0 else None
        # the grid-bounded packet walk takes its packets heaviest first
        perm = self.query_order(p, by_cost=sg is not None and not flags & (self.UDF_LANE | self.UDF_SPLIT)) if order and K > 0 else None
        with torch.cuda.device(self.device):
            st = _lib.stream(self.device)
            if sg is not None:
                grid, gface, lo, h = sg
                _lib.call("a3d_bvh_udf_grid", _lib.ptr(self.blob), _lib.ptr(p), K, _lib.ptr(perm), _lib.ptr(grid),
                          _lib.ptr(gface), _lib.ptr(self.vertices), _lib.ptr(self.faces), self.SEED_GRID_RES,
                          (ctypes.c_float * 3)(*lo), h, h / 1024, _lib.ptr(dist), _lib.ptr(face), _lib.ptr(closest),
                          _lib.ptr(uvw), int(flags), st)
            else:
                _lib.call(
                    "a3d_bvh_udf", _lib.ptr(self.blob), _lib.ptr(p), K, _lib.ptr(perm), _lib.ptr(dist),
                    _lib.ptr(face), _lib.ptr(closest), _lib.ptr(uvw), int(flags), _lib.ptr(counters), st,
                )
        return dist, face, closest, uvw

    # ------------------------------------------------------------ rays
    def _rays(self, rays_o: torch.Tensor, rays_d: torch.Tensor):
        o, d = self._points(rays_o), self._points(rays_d)
        if o.shape != d.shape:
            raise ValueError(f"rays_o {tuple(o.shape)} and rays_d {tuple(d.shape)} differ")
        return o, d

    def ray_nearest(self, rays_o: torch.Tensor, rays_d: torch.Tensor, max_t: float = 1e10):
        """(hit [K] bool, t [K] (max_t when missed), face_ids [K] int32 (-1),
        hit_points [K, 3] (the origin when missed), uvw [K, 3] (zeros when
        missed)): the nearest face, minimum t over all faces, ties to the
        lowest face index."""
        o, d = self._rays(rays_o, rays_d)
        K = o.shape[0]
        dev = self.device
        hit = torch.empty(K, dtype=torch.uint8, device=dev)
        t = torch.empty(K, dtype=torch.float32, device=dev)
        f = torch.empty(K, dtype=torch.int32, device=dev)
        p = torch.empty(K, 3, dtype=torch.float32, device=dev)
        w = torch.empty(K, 3, dtype=torch.float32, device=dev)
        with torch.cuda.device(dev):
            _lib.call("a3d_ray_nearest", _lib.ptr(self.blob), _lib.ptr(o), _lib.ptr(d), K, float(max_t), _lib.ptr(hit),
                      _lib.ptr(t), _lib.ptr(f), _lib.ptr(p), _lib.ptr(w), _lib.stream(dev))
        return hit.view(torch.bool), t, f, p, w

    def ray_intersect(self, rays_o: torch.Tensor, rays_d: torch.Tensor, max_t: float = 1e10):
        """(hit_mask, hit_t, face_ids, hit_points), as the reference's
        BVHAccelerator.ray_intersect (kernels/bvh.py:119-147)."""
        return self.ray_nearest(rays_o, rays_d, max_t)[:4]

    def _all_hits(self, o: torch.Tensor, d: torch.Tensor, max_t: float, with_ts: bool):
        """Walk-order hit lists: (offsets [K+1] int64, ray_idx, face[, t, sign])."""
        K = o.shape[0]
        dev = self.device
        with torch.cuda.device(dev):
            s = _lib.stream(dev)
            counts = torch.empty(K, dtype=torch.int32, device=dev)
            _lib.call("a3d_ray_all_count", _lib.ptr(self.blob), _lib.ptr(o), _lib.ptr(d), K, float(max_t),
                      _lib.ptr(counts), s)
            offsets = torch.zeros(K + 1, dtype=torch.int64, device=dev)
            torch.cumsum(counts, 0, dtype=torch.int64, out=offsets[1:])
            H = int(offsets[-1]) if K else 0
            ray_idx = torch.empty(H, dtype=torch.int32, device=dev)
            face = torch.empty(H, dtype=torch.int32, device=dev)
            t = torch.empty(H, dtype=torch.float32, device=dev) if with_ts else None
            sign = torch.empty(H, dtype=torch.int32, device=dev) if with_ts else None
            if H:
                _lib.call("a3d_ray_all_hits", _lib.ptr(self.blob), _lib.ptr(o), _lib.ptr(d), K, float(max_t),
                          _lib.ptr(offsets), _lib.ptr(ray_idx), _lib.ptr(face), _lib.ptr(t), _lib.ptr(sign), s)
        return offsets, ray_idx, face, t, sign

    def ray_all_hit(self, rays_o: torch.Tensor, rays_d: torch.Tensor, max_t: float = 1e10):
        """Every hit of every ray (the reference's filters: 0 < t < max_t, not
        grazing): (ray_idx, t, face_id, sign, hit_points, uvw), sorted by (ray,
        face) (the reference lists them in its tree's walk order)."""
        o, d = self._rays(rays_o, rays_d)
        dev = self.device
        _, ray_idx, face, _, _ = self._all_hits(o, d, max_t, with_ts=False)
        H = ray_idx.shape[0]
        t = torch.empty(H, dtype=torch.float32, device=dev)
        sign = torch.empty(H, dtype=torch.int32, device=dev)
        pts = torch.empty(H, 3, dtype=torch.float32, device=dev)
        uvw = torch.empty(H, 3, dtype=torch.float32, device=dev)
        if H:
            with torch.cuda.device(dev):
                s = _lib.stream(dev)
                wsb = ctypes.c_int64()
                _lib.call("a3d_pairs_sort_ws_bytes", H, ctypes.byref(wsb))
                ws = torch.empty(wsb.value, dtype=torch.uint8, device=dev)
                _lib.call("a3d_pairs_sort", _lib.ptr(ray_idx), _lib.ptr(face), H, _lib.ptr(ws), wsb.value, s)
                _lib.call("a3d_ray_eval", _lib.ptr(self.vertices), self.vertices.shape[0], _lib.ptr(self.faces),
                          self.num_faces, _lib.ptr(o), _lib.ptr(d), H, _lib.ptr(ray_idx), _lib.ptr(face),
                          float(max_t), _lib.ptr(t), _lib.ptr(sign), _lib.ptr(pts), _lib.ptr(uvw), s)
        return ray_idx, t, face, sign, pts, uvw

    def signed_crossings(self, rays_o: torch.Tensor, rays_d: torch.Tensor, max_t: float = 1e6,
                         coalesce_t: float = 1e-4) -> torch.Tensor:
        """Signed crossing count per ray (int32 [K]): hits sorted by t,
        near-coincident hits (closer than coalesce_t) merged, a group of one
        sign counting once, a mixed group by the parity of its sign sum."""
        o, d = self._rays(rays_o, rays_d)
        K = o.shape[0]
        out = torch.zeros(K, dtype=torch.int32, device=self.device)
        if K == 0:
            return out
        offsets, _, _, t, sign = self._all_hits(o, d, max_t, with_ts=True)
        with torch.cuda.device(self.device):
            _lib.call("a3d_ray_crossings", _lib.ptr(t), _lib.ptr(sign), _lib.ptr(offsets), K, float(coalesce_t),
                      _lib.ptr(out), _lib.stream(self.device))
        return out

    CROSS_OVERFLOW = -(2 ** 31)

    def signed_crossings_dirs(self, points: torch.Tensor, dirs: torch.Tensor, max_t: float = 1e6,
                              coalesce_t: float = 1e-4) -> torch.Tensor:
        """Signed crossing counts (int32 [n_dirs, N]) of every point along
        every direction, as signed_crossings gives them ray by ray: one walk
        per (point, direction) with its hits sorted in registers; the rays
        with more hits than that holds (a3d_ray_crossings_dirs marks them)
        are counted again by the two-pass path."""
        p = self._points(points)
        dv = self._points(dirs)
        N, n_dirs = p.shape[0], dv.shape[0]
        out = torch.zeros((n_dirs, N), dtype=torch.int32, device=self.device)
        if N == 0 or n_dirs == 0:
            return out
        with torch.cuda.device(self.device):
            _lib.call("a3d_ray_crossings_dirs", _lib.ptr(self.blob), _lib.ptr(p), N, _lib.ptr(dv), n_dirs,
                      float(max_t), float(coalesce_t), _lib.ptr(out), _lib.stream(self.device))
        ovf = torch.nonzero(out.reshape(-1) == self.CROSS_OVERFLOW).reshape(-1)
        if ovf.numel():
            o = p[ovf % N].contiguous()
            d = dv[ovf // N].contiguous()
            out.view(-1)[ovf] = self.signed_crossings(o, d, max_t=max_t, coalesce_t=coalesce_t)
        return out

    def aabb_hit(self, query_min: torch.Tensor, query_max: torch.Tensor) -> torch.Tensor:
        """Hit mask only (stops at the first intersecting face per box)."""
        lo, hi = self._points(query_min), self._points(query_max)
        K = lo.shape[0]
        hit = torch.empty(K, dtype=torch.uint8, device=self.device)
        with torch.cuda.device(self.device):
            _lib.call("a3d_bvh_aabb_hit", _lib.ptr(self.box_blob), _lib.ptr(lo), _lib.ptr(hi), K, _lib.ptr(hit),
                      _lib.stream(self.device))
        return hit.view(torch.bool)

    def aabb_intersect(self, query_min: torch.Tensor, query_max: torch.Tensor):
        """(hit_mask [K] bool, aabb_ids [P] int32, face_ids [P] int32), pairs
        sorted by (aabb_id, face_id)."""
        lo, hi = self._points(query_min), self._points(query_max)
        K = lo.shape[0]
        dev = self.device
        with torch.cuda.device(dev):
            counts = torch.empty(K, dtype=torch.int32, device=dev)
            s = _lib.stream(dev)
            _lib.call("a3d_bvh_aabb_count", _lib.ptr(self.box_blob), _lib.ptr(lo), _lib.ptr(hi), K, _lib.ptr(counts),
                      s)
            hit = counts > 0
            if K == 0:
                e = torch.empty(0, dtype=torch.int32, device=dev)
                return hit, e, e.clone()
            incl = torch.cumsum(counts, 0, dtype=torch.int64)
            total = int(incl[-1])
            offsets = incl - counts
            pa = torch.empty(total, dtype=torch.int32, device=dev)
            pf = torch.empty(total, dtype=torch.int32, device=dev)
            if total > 0:
                _lib.call("a3d_bvh_aabb_pairs", _lib.ptr(self.box_blob), _lib.ptr(lo), _lib.ptr(hi), K,
                          _lib.ptr(offsets), _lib.ptr(pa), _lib.ptr(pf), s)
                wsb = ctypes.c_int64()
                _lib.call("a3d_pairs_sort_ws_bytes", total, ctypes.byref(wsb))
                ws = torch.empty(wsb.value, dtype=torch.uint8, device=dev)
                _lib.call("a3d_pairs_sort", _lib.ptr(pa), _lib.ptr(pf), total, _lib.ptr(ws), wsb.value, s)
        return hit, pa, pf

    def block_cells(self, b0, cell, origin, dims) -> torch.Tensor:
        """int64 [K,3]: the cells of the grid block origin + [0, dims) whose
        box (float(ijk) * cell + b0, + cell) intersects a face by the SAT
        test, in meshgrid('ij') order (host float[3] b0, cell)."""
        dev = self.device
        n = int(dims[0]) * int(dims[1]) * int(dims[2])
        nb = ctypes.c_int64()
        _lib.call("a3d_vox_ws_bytes", n, ctypes.byref(nb))
        f3 = ctypes.c_float * 3
        i3 = ctypes.c_int64 * 3
        b0c, cc = f3(*[float(x) for x in b0]), f3(*[float(x) for x in cell])
        oc, dc = i3(*[int(x) for x in origin]), i3(*[int(x) for x in dims])
        with torch.cuda.device(dev):
            ws = torch.empty(max(nb.value, 1), dtype=torch.uint8, device=dev)
            nh = ctypes.c_int64()
            s = _lib.stream(dev)
            _lib.call("a3d_vox_cells", _lib.ptr(self.box_blob), b0c, cc, oc, dc, _lib.ptr(ws), nb.value,
                      ctypes.byref(nh), s)
            out = torch.empty((nh.value, 3), dtype=torch.int64, device=dev)
            if nh.value:
                _lib.call("a3d_vox_cells_get", _lib.ptr(ws), oc, dc, nh.value, _lib.ptr(out), s)
        return out

    def cells_sat_pairs(self, ijk: torch.Tensor, b0, cell):
        """(counts [K] int32, cell rows [P] int32, face ids [P] int32) of the
        SAT pairs of cells ijk [K,3] (int64), pairs in (cell, face) order."""
        dev = self.device
        ijk = ijk.to(dev, torch.int64).contiguous()
        K = int(ijk.shape[0])
        f3 = ctypes.c_float * 3
        b0c, cc = f3(*[float(x) for x in b0]), f3(*[float(x) for x in cell])
        with torch.cuda.device(dev):
            s = _lib.stream(dev)
            counts = torch.empty(K, dtype=torch.int32, device=dev)
            if K == 0:
                e = torch.empty(0, dtype=torch.int32, device=dev)
                return counts, e, e.clone()
            _lib.call("a3d_cells_sat_count", _lib.ptr(self.box_blob), _lib.ptr(ijk), K, b0c, cc, _lib.ptr(counts),
                      s)
            incl = torch.cumsum(counts, 0, dtype=torch.int64)
            total = int(incl[-1])
            offsets = incl - counts
            pa = torch.empty(total, dtype=torch.int32, device=dev)
            pf = torch.empty(total, dtype=torch.int32, device=dev)
            if total > 0:
                _lib.call("a3d_cells_sat_pairs", _lib.ptr(self.box_blob), _lib.ptr(ijk), K, b0c, cc,
                          _lib.ptr(offsets),
                          _lib.ptr(pa), _lib.ptr(pf), s)
                wsb = ctypes.c_int64()
                _lib.call("a3d_pairs_sort_ws_bytes", total, ctypes.byref(wsb))
                ws = torch.empty(wsb.value, dtype=torch.uint8, device=dev)
                _lib.call("a3d_pairs_sort", _lib.ptr(pa), _lib.ptr(pf), total, _lib.ptr(ws), wsb.value, s)
        return counts, pa, pf


def bvh_available() -> bool:
    """True when the HIP library loads and a GPU is visible."""
    try:
        _lib.lib()
    except Exception:
        return False
    return torch.cuda.is_available()
