// Voxel-to-dual-graph glue for gfx950: the integer formats between the flood
// fill / octree and the SDF / DMC steps.
//
// Replaces, with the reference's outputs and order:
//   - construct_surface_dual_graph (Atom3D atom3d/apps/sparse_flood_fill.py:
//     230-291): unique active cells, their unique corner vertices, the cube ->
//     vertex table and the 6-neighbour cell table;
//   - CubeGrid.voxel_unique_vertices (atom3d/grid/cube_grid.py:237-260);
//   - get_dam_wet_faces / _edges / _corners (sparse_flood_fill.py:169-227).
// The reference builds these with torch.unique(dim=0) (a lexicographic sort),
// searchsorted over flat keys, and a dense resolution^3 bool mask.  Here, up
// to 2048^3 (and when the grid is not too sparse), the cells and the corner
// vertices are bitmaps ranked by a popcount scan: unique lists are the set
// bits in order and every lookup is a rank (method 2).  Otherwise the unique
// steps are one radix sort of flat keys (lexicographic order for coordinates
// in range) with head flags and a scan, the cube -> vertex ranks come from the
// same scan, neighbours from binary search (method 1).  The collision set of
// the wet masks is a bitmap of resolution^3 bits.
#include "a3d_prim.hpp"

#include "a3d_common.hpp"

namespace a3d {
namespace {

constexpr int kBlk = 256;

inline unsigned gridn(int64_t n) { return (unsigned)div_up(n > 0 ? n : 1, kBlk); }
inline int64_t al256(int64_t x) { return (x + 255) & ~int64_t(255); }

// 6-neighbour offsets +x, -x, +y, -y, +z, -z (constants.py:47-52)
__constant__ int8_t kN6[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
// Per-edge neighbour offsets in the reference's edge numbering
// (constants.py:73-93; the table as given there).
__constant__ int8_t kEdgeNb[12][3][3] = {
    {{0, 1, 0}, {0, 0, 1}, {0, 1, 1}},     {{0, -1, 0}, {0, 0, 1}, {0, -1, 1}},
    {{0, 1, 0}, {0, 0, -1}, {0, 1, -1}},   {{0, -1, 0}, {0, 0, -1}, {0, -1, -1}},
    {{1, 0, 0}, {0, 0, 1}, {1, 0, 1}},     {{-1, 0, 0}, {0, 0, 1}, {-1, 0, 1}},
    {{1, 0, 0}, {0, 0, -1}, {1, 0, -1}},   {{-1, 0, 0}, {0, 0, -1}, {-1, 0, -1}},
    {{1, 0, 0}, {0, 1, 0}, {1, 1, 0}},     {{-1, 0, 0}, {0, 1, 0}, {-1, 1, 0}},
    {{1, 0, 0}, {0, -1, 0}, {1, -1, 0}},   {{-1, 0, 0}, {0, -1, 0}, {-1, -1, 0}}};

// Workspace for up to n input cells (8 n vertex candidates).
struct CellWs {
  void* kin;    // [8n] keys (K)
  void* kout;   // [8n]
  int32_t* vin;   // [8n] slot
  int32_t* vout;  // [8n]
  int32_t* flag;  // [8n + 1]
  int32_t* rank;  // [8n + 1]
  int64_t* count; // [2]: 0 items, 1 out-of-range inputs
  void* cub;
  size_t cub_bytes;
  int64_t total;
};

template <typename K>
size_t cell_cub_need(int64_t n8) {
  size_t b = 0, t = 0;
  (void)rocprim::radix_sort_pairs(nullptr, t, (K*)nullptr, (K*)nullptr, (int32_t*)nullptr,
                                           (int32_t*)nullptr, (int)n8, 0, (int)(8 * sizeof(K)));
  b = t > b ? t : b;
  (void)a3d::exclusive_sum(nullptr, t, (int32_t*)nullptr, (int32_t*)nullptr, (int)(n8 + 1));
  b = t > b ? t : b;
  return b;
}

CellWs carve_cells(void* base, int64_t n) {
  CellWs w;
  const int64_t n8 = 8 * (n > 1 ? n : 1);
  char* b = (char*)base;
  int64_t o = 0;
  auto take = [&](int64_t bytes) { char* p = b ? b + o : nullptr; o += al256(bytes); return p; };
  w.kin = take(8 * n8);
  w.kout = take(8 * n8);
  w.vin = (int32_t*)take(4 * n8);
  w.vout = (int32_t*)take(4 * n8);
  w.flag = (int32_t*)take(4 * (n8 + 1));
  w.rank = (int32_t*)take(4 * (n8 + 1));
  w.count = (int64_t*)take(16);
  size_t c32 = cell_cub_need<uint32_t>(n8), c64 = cell_cub_need<uint64_t>(n8);
  w.cub_bytes = c32 > c64 ? c32 : c64;
  w.cub = take((int64_t)w.cub_bytes);
  w.total = o;
  return w;
}

int bits_for(uint64_t maxval) {
  int b = 1;
  while (b < 64 && (maxval >> b) != 0) ++b;
  return b;
}

// Flat key of cell n (corner 0) or of its corner k in a grid of side P:
// (i * P + j) * P + k, lexicographic for coordinates in [0, P).
template <typename K>
__global__ void k_cell_keys(const int32_t* __restrict__ ijk, int64_t n, int corners, int64_t R, int64_t P,
                            K* __restrict__ key, int32_t* __restrict__ val, int64_t* __restrict__ count) {
  int64_t t = blockIdx.x * (int64_t)kBlk + threadIdx.x;
  if (t >= n * corners) return;
  const int64_t c = t / corners;
  const int k = (int)(t % corners);
  int64_t x = ijk[c * 3], y = ijk[c * 3 + 1], z = ijk[c * 3 + 2];
  if (x < 0 || y < 0 || z < 0 || x >= R || y >= R || z >= R) {
    atomicAdd((unsigned long long*)&count[1], 1ull);
    x = y = z = 0;
  }
  x += (k >> 0) & 1;
  y += (k >> 1) & 1;
  z += (k >> 2) & 1;
  key[t] = (K)((x * P + y) * P + z);
  val[t] = (int32_t)t;
}

template <typename K>
__global__ void k_heads(const K* __restrict__ key, int64_t n, int32_t* __restrict__ flag) {
  int64_t j = blockIdx.x * (int64_t)kBlk + threadIdx.x;
  if (j <= n) flag[j] = (j < n && (j == 0 || key[j - 1] != key[j])) ? 1 : 0;
}

// Unique keys compacted in place order; the slot's rank scattered to its
// original position (the cube -> vertex table).
template <typename K>
__global__ void k_rank_scatter(const K* __restrict__ key, const int32_t* __restrict__ slot,
                               const int32_t* __restrict__ flag, const int32_t* __restrict__ rank, int64_t n,
                               K* __restrict__ ukey, int64_t* __restrict__ slot_rank, int64_t* __restrict__ count) {
  int64_t j = blockIdx.x * (int64_t)kBlk + threadIdx.x;
  if (j == 0) count[0] = rank[n];
  if (j >= n) return;
  const int32_t r = rank[j] + (flag[j] ? 0 : -1);
  if (flag[j]) ukey[r] = key[j];
  if (slot_rank) slot_rank[slot[j]] = r;
}

template <typename K>
__global__ void k_unravel_cells(const K* __restrict__ ukey, const int64_t* __restrict__ nu, int64_t P,
                                int32_t* __restrict__ ijk) {
  int64_t u = blockIdx.x * (int64_t)kBlk + threadIdx.x;
  if (u >= *nu) return;
  const int64_t k = (int64_t)ukey[u];
  ijk[u * 3 + 0] = (int32_t)(k / (P * P));
  ijk[u * 3 + 1] = (int32_t)((k / P) % P);
  ijk[u * 3 + 2] = (int32_t)(k % P);
}

// Vertex outputs: flat index in the (R+1)^3 vertex grid, ijk, and the world
// position lo + ijk * cell as two float32 roundings (ijk.float() * cell_size
// then + bounds[0], cube_grid.py:166-167 / sparse_flood_fill.py:287-289).
template <typename K>
__global__ void k_vertex_out(const K* __restrict__ ukey, int64_t V, int64_t P, int64_t* __restrict__ vkey,
                             int64_t* __restrict__ vijk, float* __restrict__ world, float lx, float ly, float lz,
                             float cx, float cy, float cz) {
  int64_t u = blockIdx.x * (int64_t)kBlk + threadIdx.x;
  if (u >= V) return;
  const int64_t k = (int64_t)ukey[u];
  const int64_t a = k / (P * P), b = (k / P) % P, c = k % P;
  if (vkey) vkey[u] = k;
  if (vijk) {
    vijk[u * 3 + 0] = a;
    vijk[u * 3 + 1] = b;
    vijk[u * 3 + 2] = c;
  }
  if (world) {
    world[u * 3 + 0] = (float)a * cx + lx;
    world[u * 3 + 1] = (float)b * cy + ly;
    world[u * 3 + 2] = (float)c * cz + lz;
  }
}

// 6-neighbour table: the neighbour's flat key (i+di) R^2 + (j+dj) R + (k+dk),
// computed without a range check as the reference does (so a step off one
// face of the grid lands on the flat key of a cell in the next row), found
// by binary search over the sorted unique cells, else -1
// (sparse_flood_fill.py:273-283).
__global__ void k_neighbors(const int32_t* __restrict__ cells, int64_t n, int64_t R, int64_t* __restrict__ adj) {
  int64_t t = blockIdx.x * (int64_t)kBlk + threadIdx.x;
  if (t >= 6 * n) return;
  const int64_t c = t / 6;
  const int d = (int)(t % 6);
  const int64_t want = ((int64_t)cells[c * 3] + kN6[d][0]) * R * R + ((int64_t)cells[c * 3 + 1] + kN6[d][1]) * R +
                       ((int64_t)cells[c * 3 + 2] + kN6[d][2]);
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    const int64_t km = ((int64_t)cells[mid * 3] * R + cells[mid * 3 + 1]) * R + cells[mid * 3 + 2];
    if (km < want) lo = mid + 1; else hi = mid;
  }
  int64_t r = -1;
  if (lo < n) {
    const int64_t kl = ((int64_t)cells[lo * 3] * R + cells[lo * 3 + 1]) * R + cells[lo * 3 + 2];
    if (kl == want) r = lo;
  }
  adj[t] = r;
}

// ------------------------------------------------------------------ wet masks
__global__ void k_fill_u32(uint32_t* __restrict__ p, int64_t n) {
  int64_t t = blockIdx.x * (int64_t)kBlk + threadIdx.x;
  if (t < n) p[t] = 0u;
}

__global__ void k_mark(const int32_t* __restrict__ ijk, int64_t n, int64_t R, uint32_t* __restrict__ bits) {
  int64_t t = blockIdx.x * (int64_t)kBlk + threadIdx.x;
  if (t >= n) return;
  const int64_t x = ijk[t * 3], y = ijk[t * 3 + 1], z = ijk[t * 3 + 2];
  if (x < 0 || y < 0 || z < 0 || x >= R || y >= R || z >= R) return;
  const int64_t lin = (x * R + y) * R + z;
  atomicOr(&bits[lin >> 5], 1u << (lin & 31));
}

__device__ __forceinline__ bool wet_at(const uint32_t* __restrict__ bits, int64_t R, int64_t x, int64_t y,
                                       int64_t z) {
  if (x < 0 || y < 0 || z < 0 || x >= R || y >= R || z >= R) return false;
  const int64_t lin = (x * R + y) * R + z;
  return (bits[lin >> 5] >> (lin & 31)) & 1u;
}

// kind 0: faces [N,6] (neighbour d wet); 1: edges [N,12] (any of the edge's
// 3 neighbours wet); 2: corners [N,8] (any of the corner's 7 neighbours
// wet: sign offsets s = 2 * corner bit - 1 on x, y, z, xy, xz, yz, xyz).
__global__ void k_wet(int kind, const int32_t* __restrict__ dam, int64_t n, int64_t R,
                      const uint32_t* __restrict__ bits, uint8_t* __restrict__ out) {
  const int per = kind == 0 ? 6 : (kind == 1 ? 12 : 8);
  int64_t t = blockIdx.x * (int64_t)kBlk + threadIdx.x;
  if (t >= n * per) return;
  const int64_t c = t / per;
  const int q = (int)(t % per);
  const int64_t x = dam[c * 3], y = dam[c * 3 + 1], z = dam[c * 3 + 2];
  bool w = false;
  if (kind == 0) {
    w = wet_at(bits, R, x + kN6[q][0], y + kN6[q][1], z + kN6[q][2]);
  } else if (kind == 1) {
    for (int m = 0; m < 3; ++m) w = w || wet_at(bits, R, x + kEdgeNb[q][m][0], y + kEdgeNb[q][m][1], z + kEdgeNb[q][m][2]);
  } else {
    const int sx = 2 * ((q >> 0) & 1) - 1, sy = 2 * ((q >> 1) & 1) - 1, sz = 2 * ((q >> 2) & 1) - 1;
    for (int m = 1; m < 8; ++m)
      w = w || wet_at(bits, R, x + ((m & 1) ? sx : 0), y + ((m & 2) ? sy : 0), z + ((m & 4) ? sz : 0));
  }
  out[t] = w ? 1 : 0;
}

// ------------------------------------------------------------------ bitmaps
// Grids up to 2048^3: the cell set as a bitmap of R x R rows of Pzc =
// round_up(R, 32) bits (bit k of row (i, j) = cell (i, j, k)); the corner
// vertices as an (R+1) x (R+1) x Pzv bitmap derived from it with word-wide
// ORs of shifted rows; each bitmap ranked by an exclusive scan of its words'
// popcounts.  Sorted unique lists are the set bits in order, a rank lookup is
// one load of the word, one of its prefix and a popcount.
struct BmGeom {
  int64_t R, wc, wv;  // grid side; words per cell row / vertex row
  __host__ __device__ int64_t cell_words() const { return R * R * wc; }
  __host__ __device__ int64_t vert_words() const { return (R + 1) * (R + 1) * wv; }
};

inline BmGeom bm_geom(int64_t R) {
  BmGeom g;
  g.R = R;
  g.wc = (R + 31) / 32;
  g.wv = (R + 1 + 31) / 32;
  return g;
}

__global__ void k_bm_mark(const int32_t* __restrict__ ijk, int64_t n, BmGeom g, uint32_t* __restrict__ bits,
                          int64_t* __restrict__ count) {
  int64_t t = blockIdx.x * (int64_t)kBlk + threadIdx.x;
  if (t >= n) return;
  const int64_t x = ijk[t * 3], y = ijk[t * 3 + 1], z = ijk[t * 3 + 2];
  if (x < 0 || y < 0 || z < 0 || x >= g.R || y >= g.R || z >= g.R) {
    atomicAdd((unsigned long long*)&count[1], 1ull);
    return;
  }
  atomicOr(&bits[(x * g.R + y) * g.wc + (z >> 5)], 1u << (z & 31));
}

// Vertex (a, b, c) exists when a cell (a - dx, b - dy, c - dz), dx, dy, dz in
// {0, 1}, does: per vertex word, the OR over the 4 (dx, dy) rows of the cell
// word and the cell word shifted up by one bit (carry from the word below).
__global__ void k_bm_vertices(const uint32_t* __restrict__ cbits, BmGeom g, uint32_t* __restrict__ vbits) {
  int64_t t = blockIdx.x * (int64_t)kBlk + threadIdx.x;
  if (t >= g.vert_words()) return;
  const int64_t row = t / g.wv, w = t % g.wv;
  const int64_t a = row / (g.R + 1), b = row % (g.R + 1);
  uint32_t v = 0;
  for (int d = 0; d < 4; ++d) {
    const int64_t x = a - (d & 1), y = b - (d >> 1);
    if (x < 0 || y < 0 || x >= g.R || y >= g.R) continue;
    const uint32_t* r = cbits + (x * g.R + y) * g.wc;
    const uint32_t cur = w < g.wc ? r[w] : 0u;
    const uint32_t prev = (w > 0 && w - 1 < g.wc) ? r[w - 1] : 0u;
    v |= cur | (cur << 1) | (prev >> 31);
  }
  vbits[t] = v;
}

struct Popc {
  __host__ __device__ int32_t operator()(uint32_t x) const { return __builtin_popcount(x); }
};

__device__ __forceinline__ int64_t bm_rank(const uint32_t* __restrict__ bits, const int32_t* __restrict__ rank,
                                           int64_t word, int b, bool& set) {
  const uint32_t w = bits[word];
  set = (w >> b) & 1u;
  return (int64_t)rank[word] + __builtin_popcount(w & ((1u << b) - 1u));
}

// Set bits in order -> cells (int32 ijk) or vertices (flat key in the
// unpadded (R+1)^3 grid and/or int64 ijk, world position).
__global__ void k_bm_enum_cells(const uint32_t* __restrict__ bits, const int32_t* __restrict__ rank, BmGeom g,
                                int32_t* __restrict__ ijk) {
  int64_t t = blockIdx.x * (int64_t)kBlk + threadIdx.x;
  if (t >= g.cell_words()) return;
  uint32_t w = bits[t];
  if (!w) return;
  const int64_t row = t / g.wc, k0 = (t % g.wc) * 32;
  const int32_t i = (int32_t)(row / g.R), j = (int32_t)(row % g.R);
  int64_t o = rank[t];
  while (w) {
    const int b = __builtin_ctz(w);
    w &= w - 1;
    ijk[o * 3] = i;
    ijk[o * 3 + 1] = j;
    ijk[o * 3 + 2] = (int32_t)(k0 + b);
    ++o;
  }
}

__global__ void k_bm_enum_verts(const uint32_t* __restrict__ bits, const int32_t* __restrict__ rank, BmGeom g,
                                int64_t* __restrict__ vkey, int64_t* __restrict__ vijk, float* __restrict__ world,
                                float lx, float ly, float lz, float cx, float cy, float cz) {
  int64_t t = blockIdx.x * (int64_t)kBlk + threadIdx.x;
  if (t >= g.vert_words()) return;
  uint32_t w = bits[t];
  if (!w) return;
  const int64_t P = g.R + 1;
  const int64_t row = t / g.wv, c0 = (t % g.wv) * 32;
  const int64_t a = row / P, b = row % P;
  int64_t o = rank[t];
  while (w) {
    const int bit = __builtin_ctz(w);
    w &= w - 1;
    const int64_t c = c0 + bit;
    if (vkey) vkey[o] = (a * P + b) * P + c;
    if (vijk) {
      vijk[o * 3] = a;
      vijk[o * 3 + 1] = b;
      vijk[o * 3 + 2] = c;
    }
    if (world) {
      world[o * 3] = (float)a * cx + lx;
      world[o * 3 + 1] = (float)b * cy + ly;
      world[o * 3 + 2] = (float)c * cz + lz;
    }
    ++o;
  }
}

// cube_vert[n, k] = rank of corner k of cell n among the vertices.
__global__ void k_bm_cube_vert(const int32_t* __restrict__ ijk, int64_t n, BmGeom g, const uint32_t* __restrict__ vbits,
                               const int32_t* __restrict__ vrank, int64_t* __restrict__ out) {
  int64_t t = blockIdx.x * (int64_t)kBlk + threadIdx.x;
  if (t >= 8 * n) return;
  const int64_t c = t >> 3;
  const int k = (int)(t & 7);
  const int64_t x = ijk[c * 3] + (k & 1), y = ijk[c * 3 + 1] + ((k >> 1) & 1), z = ijk[c * 3 + 2] + ((k >> 2) & 1);
  bool set;
  out[t] = bm_rank(vbits, vrank, (x * (g.R + 1) + y) * g.wv + (z >> 5), (int)(z & 31), set);
}

// 6-neighbour table with the reference's unchecked flat key (see
// k_neighbors), looked up in the cell bitmap.
__global__ void k_bm_neighbors(const int32_t* __restrict__ cells, int64_t n, BmGeom g,
                               const uint32_t* __restrict__ cbits, const int32_t* __restrict__ crank,
                               int64_t* __restrict__ adj) {
  int64_t t = blockIdx.x * (int64_t)kBlk + threadIdx.x;
  if (t >= 6 * n) return;
  const int64_t c = t / 6;
  const int d = (int)(t % 6);
  const int64_t R = g.R;
  const int64_t f = ((int64_t)cells[c * 3] + kN6[d][0]) * R * R + ((int64_t)cells[c * 3 + 1] + kN6[d][1]) * R +
                    ((int64_t)cells[c * 3 + 2] + kN6[d][2]);
  int64_t r = -1;
  if (f >= 0 && f < R * R * R) {
    const int64_t x = f / (R * R), y = (f / R) % R, z = f % R;
    bool set;
    const int64_t q = bm_rank(cbits, crank, (x * R + y) * g.wc + (z >> 5), (int)(z & 31), set);
    if (set) r = q;
  }
  adj[t] = r;
}

// Sort the keys of n cells (corners = 1) or of their 8 corners (corners =
// 8) in a grid of side P, then head flags, ranks, unique keys (in kout) and,
// for corners, the per-slot rank.  Writes count[0] = unique keys.
template <typename K>
int unique_keys(const CellWs& w, const int32_t* ijk, int64_t n, int corners, int64_t R, int64_t P,
                int64_t* slot_rank, hipStream_t s) {
  const int64_t m = n * corners;
  K* kin = (K*)w.kin;
  K* kout = (K*)w.kout;
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(kBlk), 0, s, (uint32_t*)w.count, (int64_t)4);
  A3D_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_cell_keys<K>, dim3(gridn(m)), dim3(kBlk), 0, s, ijk, n, corners, R, P, kin, w.vin, w.count);
  A3D_LAUNCH_CHECK();
  const int eb = bits_for((uint64_t)(P * P * P - 1));
  size_t tb = w.cub_bytes;
  A3D_HIP_CHECK(rocprim::radix_sort_pairs(w.cub, tb, kin, kout, w.vin, w.vout, (int)m, 0, eb, s));
  hipLaunchKernelGGL(k_heads<K>, dim3(gridn(m + 1)), dim3(kBlk), 0, s, kout, m, w.flag);
  A3D_LAUNCH_CHECK();
  tb = w.cub_bytes;
  A3D_HIP_CHECK(a3d::exclusive_sum(w.cub, tb, w.flag, w.rank, (int)(m + 1), s));
  // unique keys go to kin (free after the sort)
  hipLaunchKernelGGL(k_rank_scatter<K>, dim3(gridn(m)), dim3(kBlk), 0, s, kout, w.vout, w.flag, w.rank, m, kin,
                     slot_rank, w.count);
  A3D_LAUNCH_CHECK();
  return A3D_OK;
}

int read_counts(const CellWs& w, int64_t* host, hipStream_t s, const char* who) {
  int64_t c[2];
  A3D_HIP_CHECK(hipMemcpyAsync(c, w.count, sizeof(c), hipMemcpyDeviceToHost, s));
  A3D_HIP_CHECK(hipStreamSynchronize(s));
  if (c[1] != 0) return fail(A3D_ERR_ARG, "%s: %lld coordinates outside [0, resolution)", who, (long long)c[1]);
  *host = c[0];
  return A3D_OK;
}

bool use32(int64_t P) { return P * P * P - 1 <= (int64_t)0xffffffffLL; }

// method: 0 = choose, 1 = sort, 2 = bitmap (resolution <= 2048).  The bitmap
// path is chosen when its bitmaps are not much larger than the sort arrays.
constexpr int64_t kBmMaxR = 2048;
int pick_method(int method, int64_t n, int64_t R) {
  if (method == 1 || method == 2) return method;
  const int64_t nn = n > 1 ? n : 1;
  return (R <= kBmMaxR && R * R * R <= 4096 * nn) ? 2 : 1;
}

// Bitmap workspace: cell bits and ranks, vertex bits and ranks, counters, scan
// temporary storage.
struct BmWs {
  uint32_t* cbits;
  int32_t* crank;  // [cell words + 1]
  uint32_t* vbits;
  int32_t* vrank;  // [vertex words + 1]
  int64_t* count;  // [2]
  void* cub;
  size_t cub_bytes;
  int64_t total;
};

BmWs carve_bm(void* base, const BmGeom& g) {
  BmWs w;
  char* b = (char*)base;
  int64_t o = 0;
  auto take = [&](int64_t bytes) { char* p = b ? b + o : nullptr; o += al256(bytes); return p; };
  const int64_t cw = g.cell_words(), vw = g.vert_words();
  w.cbits = (uint32_t*)take(4 * (cw + 1));  // + a zero word read by the scan
  w.crank = (int32_t*)take(4 * (cw + 1));
  w.vbits = (uint32_t*)take(4 * (vw + 1));
  w.vrank = (int32_t*)take(4 * (vw + 1));
  w.count = (int64_t*)take(16);
  size_t t = 0;
  rocprim::transform_iterator<const uint32_t*, Popc, int32_t> it(nullptr, Popc());
  (void)a3d::exclusive_sum(nullptr, t, it, (int32_t*)nullptr, (int)((cw > vw ? cw : vw) + 1));
  w.cub_bytes = t;
  w.cub = take((int64_t)t);
  w.total = o;
  return w;
}

// Ranks of `words` bitmap words (the scan runs over words + 1 items; the
// extra item reads the zero word past the end, so rank[words] = total).
int bm_scan(const BmWs& w, const uint32_t* bits, int32_t* rank, int64_t words, hipStream_t s) {
  rocprim::transform_iterator<const uint32_t*, Popc, int32_t> it(bits, Popc());
  size_t tb = w.cub_bytes;
  A3D_HIP_CHECK(a3d::exclusive_sum(w.cub, tb, it, rank, (int)(words + 1), s));
  return A3D_OK;
}

// Mark cells (input ijk, n) into the cell bitmap and rank it; vertices too
// when `verts`.  The bitmaps get one spare zero word each for the scan.
int bm_build(const BmWs& w, const BmGeom& g, const int32_t* ijk, int64_t n, bool verts, hipStream_t s) {
  const int64_t cw = g.cell_words(), vw = g.vert_words();
  hipLaunchKernelGGL(k_fill_u32, dim3(gridn(cw + 1)), dim3(kBlk), 0, s, w.cbits, cw + 1);
  A3D_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(kBlk), 0, s, (uint32_t*)w.count, (int64_t)4);
  A3D_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_bm_mark, dim3(gridn(n)), dim3(kBlk), 0, s, ijk, n, g, w.cbits, w.count);
  A3D_LAUNCH_CHECK();
  if (int rc = bm_scan(w, w.cbits, w.crank, cw, s)) return rc;
  if (verts) {
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(kBlk), 0, s, w.vbits + vw, (int64_t)1);
    A3D_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_bm_vertices, dim3(gridn(vw)), dim3(kBlk), 0, s, (const uint32_t*)w.cbits, g, w.vbits);
    A3D_LAUNCH_CHECK();
    if (int rc = bm_scan(w, w.vbits, w.vrank, vw, s)) return rc;
  }
  return A3D_OK;
}

// count[0] = rank[words], then read both counters (one host sync).
__global__ void k_bm_total(const int32_t* __restrict__ rank, int64_t words, int64_t* count) {
  if (threadIdx.x == 0 && blockIdx.x == 0) count[0] = rank[words];
}

int bm_counts(const BmWs& w, const int32_t* rank, int64_t words, int64_t* host, hipStream_t s, const char* who) {
  hipLaunchKernelGGL(k_bm_total, dim3(1), dim3(64), 0, s, rank, words, w.count);
  A3D_LAUNCH_CHECK();
  int64_t c[2];
  A3D_HIP_CHECK(hipMemcpyAsync(c, w.count, sizeof(c), hipMemcpyDeviceToHost, s));
  A3D_HIP_CHECK(hipStreamSynchronize(s));
  if (c[1] != 0) return fail(A3D_ERR_ARG, "%s: %lld coordinates outside [0, resolution)", who, (long long)c[1]);
  *host = c[0];
  return A3D_OK;
}

int check_cells_args(int64_t n, int resolution, int method, const char* who) {
  if (resolution < 1 || resolution > (1 << 20) || n < 0 || n > ((int64_t)1 << 27) || method < 0 || method > 2)
    return fail(A3D_ERR_ARG, "%s: bad arguments (n=%lld, resolution=%d, method=%d)", who, (long long)n, resolution,
                method);
  if (method == 2 && resolution > kBmMaxR)
    return fail(A3D_ERR_ARG, "%s: the bitmap method needs resolution <= %d", who, (int)kBmMaxR);
  return A3D_OK;
}

int64_t cells_ws_need(int64_t n, int64_t R, int method) {
  return pick_method(method, n, R) == 2 ? carve_bm(nullptr, bm_geom(R)).total : carve_cells(nullptr, n).total;
}

}  // namespace
}  // namespace a3d

using namespace a3d;

extern "C" int a3d_cells_ws_bytes(int64_t n, int resolution, int method, int64_t* bytes) {
  if (!bytes) return fail(A3D_ERR_ARG, "a3d_cells_ws_bytes: null output");
  if (int rc = check_cells_args(n, resolution, method, "a3d_cells_ws_bytes")) return rc;
  *bytes = cells_ws_need(n, resolution, method);
  return A3D_OK;
}

extern "C" int a3d_unique_cells(const int32_t* ijk, int64_t n, int resolution, int method, void* ws,
                                int64_t ws_bytes, int32_t* out_ijk, int64_t* n_out, a3d_stream_t stream) {
  if (!n_out) return fail(A3D_ERR_ARG, "a3d_unique_cells: null n_out");
  if (int rc = check_cells_args(n, resolution, method, "a3d_unique_cells")) return rc;
  *n_out = 0;
  if (n == 0) return A3D_OK;
  if (!ijk || !ws || !out_ijk) return fail(A3D_ERR_ARG, "a3d_unique_cells: null argument");
  const int64_t R = resolution;
  if (ws_bytes < cells_ws_need(n, R, method))
    return fail(A3D_ERR_ARG, "a3d_unique_cells: workspace %lld < %lld bytes", (long long)ws_bytes,
                (long long)cells_ws_need(n, R, method));
  hipStream_t s = (hipStream_t)stream;
  if (pick_method(method, n, R) == 2) {
    const BmGeom g = bm_geom(R);
    BmWs w = carve_bm(ws, g);
    if (int rc = bm_build(w, g, ijk, n, false, s)) return rc;
    hipLaunchKernelGGL(k_bm_enum_cells, dim3(gridn(g.cell_words())), dim3(kBlk), 0, s, (const uint32_t*)w.cbits,
                       (const int32_t*)w.crank, g, out_ijk);
    A3D_LAUNCH_CHECK();
    return bm_counts(w, w.crank, g.cell_words(), n_out, s, "a3d_unique_cells");
  }
  CellWs w = carve_cells(ws, n);
  int rc = use32(R) ? unique_keys<uint32_t>(w, ijk, n, 1, R, R, nullptr, s)
                    : unique_keys<uint64_t>(w, ijk, n, 1, R, R, nullptr, s);
  if (rc) return rc;
  if (use32(R))
    hipLaunchKernelGGL(k_unravel_cells<uint32_t>, dim3(gridn(n)), dim3(kBlk), 0, s, (const uint32_t*)w.kin,
                       (const int64_t*)w.count, R, out_ijk);
  else
    hipLaunchKernelGGL(k_unravel_cells<uint64_t>, dim3(gridn(n)), dim3(kBlk), 0, s, (const uint64_t*)w.kin,
                       (const int64_t*)w.count, R, out_ijk);
  A3D_LAUNCH_CHECK();
  return read_counts(w, n_out, s, "a3d_unique_cells");
}

extern "C" int a3d_cell_vertices(const int32_t* ijk, int64_t n, int resolution, int method, void* ws,
                                 int64_t ws_bytes, int64_t* cube_vert, int64_t* n_vert, a3d_stream_t stream) {
  if (!n_vert) return fail(A3D_ERR_ARG, "a3d_cell_vertices: null n_vert");
  if (int rc = check_cells_args(n, resolution, method, "a3d_cell_vertices")) return rc;
  *n_vert = 0;
  if (n == 0) return A3D_OK;
  if (!ijk || !ws || !cube_vert) return fail(A3D_ERR_ARG, "a3d_cell_vertices: null argument");
  const int64_t R = resolution, P = R + 1;
  if (ws_bytes < cells_ws_need(n, R, method))
    return fail(A3D_ERR_ARG, "a3d_cell_vertices: workspace %lld < %lld bytes", (long long)ws_bytes,
                (long long)cells_ws_need(n, R, method));
  hipStream_t s = (hipStream_t)stream;
  if (pick_method(method, n, R) == 2) {
    const BmGeom g = bm_geom(R);
    BmWs w = carve_bm(ws, g);
    if (int rc = bm_build(w, g, ijk, n, true, s)) return rc;
    int64_t nv = 0;
    // range errors first: out-of-range cells would index outside the bitmap
    if (int rc = bm_counts(w, w.vrank, g.vert_words(), &nv, s, "a3d_cell_vertices")) return rc;
    hipLaunchKernelGGL(k_bm_cube_vert, dim3(gridn(8 * n)), dim3(kBlk), 0, s, ijk, n, g, (const uint32_t*)w.vbits,
                       (const int32_t*)w.vrank, cube_vert);
    A3D_LAUNCH_CHECK();
    *n_vert = nv;
    return A3D_OK;
  }
  CellWs w = carve_cells(ws, n);
  int rc = use32(P) ? unique_keys<uint32_t>(w, ijk, n, 8, R, P, cube_vert, s)
                    : unique_keys<uint64_t>(w, ijk, n, 8, R, P, cube_vert, s);
  if (rc) return rc;
  return read_counts(w, n_vert, s, "a3d_cell_vertices");
}

extern "C" int a3d_cell_vertices_get(const void* ws, int64_t n, int resolution, int method, int64_t n_vert,
                                     int64_t* vkey, int64_t* vijk, float* world, const float* lo, const float* cell,
                                     a3d_stream_t stream) {
  if (!ws || n_vert < 0 || n_vert > 8 * n) return fail(A3D_ERR_ARG, "a3d_cell_vertices_get: bad arguments");
  if (int rc = check_cells_args(n, resolution, method, "a3d_cell_vertices_get")) return rc;
  if (world && (!lo || !cell)) return fail(A3D_ERR_ARG, "a3d_cell_vertices_get: world needs lo and cell");
  if (n_vert == 0) return A3D_OK;
  hipStream_t s = (hipStream_t)stream;
  const int64_t R = resolution, P = R + 1;
  const float l0 = world ? lo[0] : 0.f, l1 = world ? lo[1] : 0.f, l2 = world ? lo[2] : 0.f;
  const float c0 = world ? cell[0] : 0.f, c1 = world ? cell[1] : 0.f, c2 = world ? cell[2] : 0.f;
  if (pick_method(method, n, R) == 2) {
    const BmGeom g = bm_geom(R);
    BmWs w = carve_bm((void*)ws, g);
    hipLaunchKernelGGL(k_bm_enum_verts, dim3(gridn(g.vert_words())), dim3(kBlk), 0, s, (const uint32_t*)w.vbits,
                       (const int32_t*)w.vrank, g, vkey, vijk, world, l0, l1, l2, c0, c1, c2);
    A3D_LAUNCH_CHECK();
    return A3D_OK;
  }
  CellWs w = carve_cells((void*)ws, n);
  if (use32(P))
    hipLaunchKernelGGL(k_vertex_out<uint32_t>, dim3(gridn(n_vert)), dim3(kBlk), 0, s, (const uint32_t*)w.kin, n_vert,
                       P, vkey, vijk, world, l0, l1, l2, c0, c1, c2);
  else
    hipLaunchKernelGGL(k_vertex_out<uint64_t>, dim3(gridn(n_vert)), dim3(kBlk), 0, s, (const uint64_t*)w.kin, n_vert,
                       P, vkey, vijk, world, l0, l1, l2, c0, c1, c2);
  A3D_LAUNCH_CHECK();
  return A3D_OK;
}

extern "C" int a3d_cell_neighbors(const int32_t* cells, int64_t n, int resolution, int method, void* ws,
                                  int64_t ws_bytes, int64_t* adj, a3d_stream_t stream) {
  if (int rc = check_cells_args(n, resolution, method, "a3d_cell_neighbors")) return rc;
  if (n > 0 && (!cells || !adj)) return fail(A3D_ERR_ARG, "a3d_cell_neighbors: null argument");
  if (n == 0) return A3D_OK;
  const int64_t R = resolution;
  hipStream_t s = (hipStream_t)stream;
  if (pick_method(method, n, R) == 2) {
    if (!ws || ws_bytes < cells_ws_need(n, R, method))
      return fail(A3D_ERR_ARG, "a3d_cell_neighbors: workspace %lld < %lld bytes", (long long)ws_bytes,
                  (long long)cells_ws_need(n, R, method));
    const BmGeom g = bm_geom(R);
    BmWs w = carve_bm(ws, g);
    if (int rc = bm_build(w, g, cells, n, false, s)) return rc;
    hipLaunchKernelGGL(k_bm_neighbors, dim3(gridn(6 * n)), dim3(kBlk), 0, s, cells, n, g, (const uint32_t*)w.cbits,
                       (const int32_t*)w.crank, adj);
    A3D_LAUNCH_CHECK();
    return A3D_OK;
  }
  hipLaunchKernelGGL(k_neighbors, dim3(gridn(6 * n)), dim3(kBlk), 0, s, cells, n, R, adj);
  A3D_LAUNCH_CHECK();
  return A3D_OK;
}

extern "C" int a3d_wet_ws_bytes(int resolution, int64_t* bytes) {
  if (resolution < 1 || resolution > 4096 || !bytes)
    return fail(A3D_ERR_ARG, "a3d_wet_ws_bytes: resolution %d out of range [1, 4096]", resolution);
  const int64_t R = resolution;
  *bytes = al256(((R * R * R + 31) / 32) * 4);
  return A3D_OK;
}

extern "C" int a3d_wet_mask(int kind, const int32_t* dam, int64_t n, const int32_t* coll, int64_t c, int resolution,
                            void* bitmap, int64_t bitmap_bytes, uint8_t* out, a3d_stream_t stream) {
  if (kind < 0 || kind > 2) return fail(A3D_ERR_ARG, "a3d_wet_mask: kind %d not in {0 faces, 1 edges, 2 corners}", kind);
  if (n < 0 || c < 0 || resolution < 1 || resolution > 4096)
    return fail(A3D_ERR_ARG, "a3d_wet_mask: bad sizes (n=%lld, c=%lld, resolution=%d)", (long long)n, (long long)c,
                resolution);
  if (n == 0) return A3D_OK;
  const int64_t R = resolution, words = (R * R * R + 31) / 32;
  if (!dam || !out || !bitmap || (c > 0 && !coll)) return fail(A3D_ERR_ARG, "a3d_wet_mask: null argument");
  if (bitmap_bytes < words * 4)
    return fail(A3D_ERR_ARG, "a3d_wet_mask: bitmap %lld < %lld bytes", (long long)bitmap_bytes, (long long)(words * 4));
  hipStream_t s = (hipStream_t)stream;
  uint32_t* bits = (uint32_t*)bitmap;
  hipLaunchKernelGGL(k_fill_u32, dim3(gridn(words)), dim3(kBlk), 0, s, bits, words);
  A3D_LAUNCH_CHECK();
  if (c > 0) {
    hipLaunchKernelGGL(k_mark, dim3(gridn(c)), dim3(kBlk), 0, s, coll, c, R, bits);
    A3D_LAUNCH_CHECK();
  }
  const int per = kind == 0 ? 6 : (kind == 1 ? 12 : 8);
  hipLaunchKernelGGL(k_wet, dim3(gridn(n * per)), dim3(kBlk), 0, s, kind, dam, n, R, (const uint32_t*)bits, out);
  A3D_LAUNCH_CHECK();
  return A3D_OK;
}
