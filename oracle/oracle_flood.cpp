// Oracle — CPU restatement of the reference's dense flood fill.
//
// TEST INFRASTRUCTURE ONLY (see oracle_bvh.cpp for the ground rules).
//
// Restated from Atom3D atom3d/kernels/flood_fill_kernels.cu:21-175: labels
// -2 dry, -1 dam, 1 collision, 2 water.  Start voxel = water; every reached
// free voxel v visits its neighbours (6: L1 <= 1, 18: L1 <= 2, 26: all 26):
// an occupied neighbour becomes dam and makes v a collision voxel, a dry free
// neighbour becomes water and is reached.  The reference runs this level by
// level with atomics; its final labels do not depend on the visiting order
// (each label is a function of reachability only), so a sequential FIFO BFS
// gives the same array.
// Pinning (tests/test_oracle_flood.py): reachability against
// scipy.ndimage.label connected components of the free space, and the label
// rules against a direct neighbourhood computation with numpy.
#include <cstdint>
#include <cstdlib>
#include <vector>

extern "C" {

// occ uint8[D*H*W] (axis order d, h, w), start (z, y, x) = (d, h, w) indices;
// mask int32[D*H*W] output.  Returns the number of BFS levels.
int64_t orc_flood_fill_dense(const uint8_t* occ, int D, int H, int W, int sz, int sy, int sx, int conn,
                             int32_t* mask) {
  const int64_t n = (int64_t)D * H * W;
  for (int64_t i = 0; i < n; ++i) mask[i] = -2;
  if (sx < 0 || sx >= W || sy < 0 || sy >= H || sz < 0 || sz >= D) return 0;
  int64_t s = ((int64_t)sz * H + sy) * W + sx;
  mask[s] = 2;
  std::vector<int64_t> cur{s}, nxt;
  int64_t levels = 0;
  while (!cur.empty()) {
    ++levels;
    nxt.clear();
    for (int64_t v : cur) {
      int z = (int)(v / ((int64_t)H * W));
      int y = (int)((v / W) % H);
      int x = (int)(v % W);
      bool col = false;
      for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
          for (int dx = -1; dx <= 1; ++dx) {
            if (!dx && !dy && !dz) continue;
            int l1 = abs(dx) + abs(dy) + abs(dz);
            if (conn == 6 && l1 > 1) continue;
            if (conn == 18 && l1 > 2) continue;
            int nx = x + dx, ny = y + dy, nz = z + dz;
            if (nx < 0 || nx >= W || ny < 0 || ny >= H || nz < 0 || nz >= D) continue;
            int64_t u = ((int64_t)nz * H + ny) * W + nx;
            if (occ[u]) {
              if (mask[u] == -2) mask[u] = -1;
              col = true;
            } else if (mask[u] == -2) {
              mask[u] = 2;
              nxt.push_back(u);
            }
          }
      if (col) mask[v] = 1;
    }
    cur.swap(nxt);
  }
  return levels;
}

}  // extern "C"
