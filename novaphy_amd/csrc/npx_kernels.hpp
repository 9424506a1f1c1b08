// npx_kernels.hpp -- HIP kernels of the PBF fluid step for gfx950.
//
// Layout in HBM (one set of arrays for every scene of a context):
//   particle slots  cell-sorted; scene s owns slots [seg_off, seg_off + n),
//                   segments aligned to NPX_BLOCK so a block never straddles
//                   two scenes and its scene constants are wave-uniform.
//   X, V            float4 committed position / velocity (w unused)
//   XS[2]           float4 predicted position x* (w = lambda), ping-pong
//   RHO             float density; VT float4 (v after the speed clamp, w = m/rho~)
//   OM              float4 vorticity (xyz) and |omega| (w)
//   ID / CELL       original particle index / local window cell of x*_0
//   class tables    clsmap[window cell] -> class; cstart[class] exclusive prefix
//                   of class populations (classes of scene s contiguous).
//   skin lists      per-step candidate lists (struct Skin below)
//   runs            runs of identical state: start flags, run of each slot,
//                   first slot of each run (k_gather, k_pieces)
#pragma once

#include <type_traits>

#include "npx_device.hpp"
#include "npx_repadd.hpp"

#define NPX_EMPTY_KEY 0xFFFFFFFFFFFFFFFFull

// Global slot of this thread.  Scene segments are NPX_BLOCK-aligned, so the
// scene of a slot is blk_scene[slot / NPX_BLOCK] whatever block size a kernel
// is launched with (64..NPX_MAX_BLOCK), and a wave never straddles scenes.
__device__ __forceinline__ int npx_slot() { return (int)(blockIdx.x * blockDim.x + threadIdx.x); }

// Device error flags (bit set = condition seen).
#define NPX_FLAG_WINDOW 1     // a particle's cell fell outside the window interior
#define NPX_FLAG_NONFINITE 2  // a position or velocity is NaN / inf
// flags[0] collects every condition; flags[1 + s] those of scene s.
__device__ __forceinline__ void raise_flag(int* flags, int s, int bit) {
  atomicOr(flags, bit);
  atomicOr(flags + 1 + s, bit);
}

// ------------------------------------------------------------------ utils
__global__ void k_fill_i32(int* __restrict__ p, int v, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Exclusive scan of int32 in three phases (1024 elements per block).
__global__ void __launch_bounds__(256) k_scan_reduce(const int* __restrict__ in, int n, int* __restrict__ bsum) {
  __shared__ int red[256];
  const int base = blockIdx.x * 1024 + threadIdx.x * 4;
  int s = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (base + k < n) s += in[base + k];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) bsum[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(1024) k_scan_bsum(int* __restrict__ bsum, int nb) {
  __shared__ int tmp[1024];
  __shared__ int carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < nb; base += 1024) {
    int i = base + threadIdx.x;
    int v = (i < nb) ? bsum[i] : 0;
    tmp[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
      int t = ((int)threadIdx.x >= o) ? tmp[threadIdx.x - o] : 0;
      __syncthreads();
      tmp[threadIdx.x] += t;
      __syncthreads();
    }
    int incl = tmp[threadIdx.x];
    if (i < nb) bsum[i] = carry + incl - v;
    __syncthreads();
    if (threadIdx.x == 1023) carry += incl;
    __syncthreads();
  }
}

__global__ void __launch_bounds__(256) k_scan_down(int* __restrict__ data, int n, const int* __restrict__ bsum) {
  __shared__ int tmp[256];
  const int base = blockIdx.x * 1024 + threadIdx.x * 4;
  int v[4];
  int s = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    v[k] = (base + k < n) ? data[base + k] : 0;
    s += v[k];
  }
  tmp[threadIdx.x] = s;
  __syncthreads();
  for (int o = 1; o < 256; o <<= 1) {
    int t = ((int)threadIdx.x >= o) ? tmp[threadIdx.x - o] : 0;
    __syncthreads();
    tmp[threadIdx.x] += t;
    __syncthreads();
  }
  int run = bsum[blockIdx.x] + tmp[threadIdx.x] - s;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (base + k < n) data[base + k] = run;
    run += v[k];
  }
}

// ------------------------------------------------------ class tables
// Cells of a window whose Teschner keys are equal share one class (the
// reference's hash bucket, neighbor_search.cpp:14-22); class = smallest local
// cell index carrying that key.  Built once per window with an open-addressing
// table; atomicMin makes the result independent of insertion order.
__global__ void k_ht_clear(unsigned long long* __restrict__ tk, int* __restrict__ tv, int cap) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < cap) {
    tk[i] = NPX_EMPTY_KEY;
    tv[i] = 0x7FFFFFFF;
  }
}

__device__ __forceinline__ unsigned ht_slot(uint64_t key, unsigned mask) {
  return (unsigned)((key * 0x9E3779B97F4A7C15ull) >> 32) & mask;
}

__global__ void k_ht_insert(unsigned long long* __restrict__ tk, int* __restrict__ tv, unsigned mask, int wlo0,
                            int wlo1, int wlo2, int wx, int wy, int ncell) {
  int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= ncell) return;
  int lx = l % wx, t = l / wx, ly = t % wy, lz = t / wy;
  uint64_t key = teschner_key(wlo0 + lx, wlo1 + ly, wlo2 + lz);
  if (key == NPX_EMPTY_KEY) return;  // cannot be stored; such a cell is its own class
  unsigned s = ht_slot(key, mask);
  for (unsigned probe = 0; probe <= mask; ++probe) {
    unsigned long long prev = atomicCAS(&tk[s], NPX_EMPTY_KEY, (unsigned long long)key);
    if (prev == NPX_EMPTY_KEY || prev == key) {
      atomicMin(&tv[s], l);
      return;
    }
    s = (s + 1) & mask;
  }
}

__global__ void k_ht_lookup(const unsigned long long* __restrict__ tk, const int* __restrict__ tv, unsigned mask,
                            int wlo0, int wlo1, int wlo2, int wx, int wy, int ncell, int cls_base,
                            int* __restrict__ clsmap) {
  int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= ncell) return;
  int lx = l % wx, t = l / wx, ly = t % wy, lz = t / wy;
  uint64_t key = teschner_key(wlo0 + lx, wlo1 + ly, wlo2 + lz);
  int cls = l;
  if (key != NPX_EMPTY_KEY) {
    unsigned s = ht_slot(key, mask);
    for (unsigned probe = 0; probe <= mask; ++probe) {
      if (tk[s] == key) {
        cls = tv[s];
        break;
      }
      s = (s + 1) & mask;
    }
  }
  clsmap[l] = cls_base + cls;
}

// ------------------------------------------------------ fluid sort
// Class counting of the binning kernels (k_predict_bin, k_aux_bin,
// k_bin_scene_points), blocks of NPX_BIN_BLOCK threads.  Points arrive in an
// order where equal classes sit together (the previous step's class order),
// so a wave holds a few segments of one class each.  Each segment's first
// lane adds the segment's length to its class's entry in a per-block LDS table
// (open addressing, room for every thread's class), and once the block's
// table is complete each entry takes one global atomic.  A stack of many
// particles in one cell thus costs one global atomic per block instead of one
// per lane or per wave, and a wave's segments count in parallel.  The offsets
// only place points provisionally; the rank kernels fix the order.
#define NPX_BIN_BLOCK 1024
#define NPX_BIN_TABLE 2048

struct BinTable {
  int k[NPX_BIN_TABLE];  // class, -1: empty
  int c[NPX_BIN_TABLE];  // the block's count, then its global base
};

__device__ __forceinline__ void bin_clear(BinTable& T) {
  for (int i = (int)threadIdx.x; i < NPX_BIN_TABLE; i += (int)blockDim.x) {
    T.k[i] = -1;
    T.c[i] = 0;
  }
}

// Offset of the calling point inside class c (c >= 0 when active).  Every
// thread of the block calls it, after bin_clear.
__device__ __forceinline__ int bin_offset(BinTable& T, int c, bool active, int* __restrict__ ccount) {
  __syncthreads();  // the table is clear
  const int lane = (int)(threadIdx.x & 63);
  const int cprev = __shfl(c, max(lane - 1, 0), 64);
  const bool head = active && (lane == 0 || cprev != c);
  const unsigned long long heads = __ballot(head);
  const unsigned long long ends = heads | ~__ballot(active);  // segment starts and idle lanes
  const unsigned long long upto = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);  // lanes 0..lane
  int slot = 0, base = 0;
  if (head) {
    const unsigned long long after = ends & ~upto;
    const int len = (after ? (int)__builtin_ctzll(after) : 64) - lane;
    unsigned h = ((unsigned)c * 2654435761u) >> 21;  // 11 bits: NPX_BIN_TABLE entries
    for (;;) {  // at most blockDim distinct classes in a table of twice that: ends
      const int old = atomicCAS(&T.k[h], -1, c);
      if (old == -1 || old == c) break;
      h = (h + 1u) & (NPX_BIN_TABLE - 1);
    }
    slot = (int)h;
    base = atomicAdd(&T.c[h], len);
  }
  // an active lane's segment starts at the last head at or before it
  const int hl = active ? 63 - (int)__builtin_clzll(heads & upto) : lane;
  slot = __shfl(slot, hl, 64);
  base = __shfl(base, hl, 64);
  __syncthreads();  // the block's counts are complete
  for (int i = (int)threadIdx.x; i < NPX_BIN_TABLE; i += (int)blockDim.x) {
    const int k = T.k[i];
    if (k >= 0) T.c[i] = atomicAdd(&ccount[k], T.c[i]);
  }
  __syncthreads();
  return active ? T.c[slot] + base + (lane - hl) : 0;
}

// Pass 1 (PBFSolver::step predict, pbf_solver.cpp:28-34, and grid build
// :37-40): v' = v + g*dt, x* = x + v'*dt, cell of x*, class, and the slot
// offset inside the class (bin_offset).  Nothing is written back to X / V:
// the gather pass recomputes v' and x* with the same operations.  nblk: the
// 256-slot blocks of the slot range.
__global__ void __launch_bounds__(NPX_BIN_BLOCK)
k_predict_bin(const SceneC* __restrict__ scenes, const int* __restrict__ blk_scene, int nblk,
              const float4* __restrict__ X, const float4* __restrict__ V, const int* __restrict__ clsmap,
              int* __restrict__ ccount, int* __restrict__ cls_of, int* __restrict__ off_of, int* __restrict__ cell_of,
              int* __restrict__ flags, int* __restrict__ huge) {
  __shared__ BinTable T;
  if (blockIdx.x == 0 && threadIdx.x == 0) huge[0] = 0;  // the step's huge-class count (k_scatter)
  bin_clear(T);
  const int q = npx_slot();
  const int sb = q / NPX_BLOCK;
  const int s = (sb < nblk) ? blk_scene[sb] : -1;  // uniform per wave
  bool active = false;
  int c = -1;
  if (s >= 0) {
    const SceneC& S = scenes[s];
    active = q < S.seg_off + S.n;
    if (active) {
      const float4 x = X[q];
      const float4 v = V[q];
      const float vx = v.x + S.gdt0, vy = v.y + S.gdt1, vz = v.z + S.gdt2;
      const float sx = x.x + vx * S.dt, sy = x.y + vy * S.dt, sz = x.z + vz * S.dt;
      int lx = cell_coord(sx, S.inv_h) - S.wlo0;
      int ly = cell_coord(sy, S.inv_h) - S.wlo1;
      int lz = cell_coord(sz, S.inv_h) - S.wlo2;
      // The 27-cell neighbourhood must lie inside the window; outside it the
      // class table is undefined, so the step is flagged as failed.
      if (lx < 1 || ly < 1 || lz < 1 || lx > S.wx - 2 || ly > S.wy - 2 || lz > S.wz - 2) {
        raise_flag(flags, s, NPX_FLAG_WINDOW);
        lx = min(max(lx, 1), S.wx - 2);
        ly = min(max(ly, 1), S.wy - 2);
        lz = min(max(lz, 1), S.wz - 2);
      }
      const int l = lx + S.wx * (ly + S.wy * lz);
      c = clsmap[S.cls_base + l];
      cls_of[q] = c;
      cell_of[q] = l;
    }
  }
  const int off = bin_offset(T, c, active, ccount);
  if (active) off_of[q] = off;
}

// Classes with more particles than this are ranked through sorted blocks;
// smaller ones by direct counting.
#define NPX_RANK_DIRECT 64

// Merge launches a large class in slot range [b, e) has work in: ceil(log2)
// of the number of 256-slot blocks it spans (0 inside one block).  The merge
// launch at `level` merges the class only while this exceeds `level`.  The
// sort passes store each block's maximum (bigb) so that a merge launch's
// blocks whose classes are done exit before any per-slot load.
__device__ __forceinline__ int merge_levels(int b, int e) {
  const int np = (e - 1) / NPX_BLOCK - b / NPX_BLOCK + 1;
  return np <= 1 ? 0 : 32 - __clz(np - 1);
}

// Huge classes: those spanning more than 2^hl blocks (hl: a launch argument,
// NPX_HUGE_LEVELS by default; a scene collapsing into one cell puts most of
// its particles in one class).  Any class ranks the same either way.  Up to
// NPX_HUGE_MAX of them per step are ranked in original-index space instead
// of by merge launches (k_huge_hist / k_huge_scan / k_huge_place); further
// ones take the merge launches.  k_scatter gives each its index h (hid[c],
// -1 past the cap) and records its scene, first slot and end slot in
// huge[1 + 3h .. 3 + 3h]; huge[0] counts them.
#define NPX_HUGE_LEVELS 6
#define NPX_HUGE_MAX 64
#define NPX_HUGE_NONE 0xFF

__device__ __forceinline__ bool is_huge(int b, int e, int hl) { return merge_levels(b, e) > hl; }

// Pass 2: provisional slot = class start + atomic offset.  The point at
// offset 0 of a huge class registers it (hid, huge).
__global__ void __launch_bounds__(NPX_BLOCK)
k_scatter(const SceneC* __restrict__ scenes, const int* __restrict__ blk_scene, const int* __restrict__ cstart,
          const int* __restrict__ cls_of, const int* __restrict__ off_of, const int* __restrict__ id,
          int* __restrict__ tperm, int* __restrict__ tkey, int* __restrict__ hid, int* __restrict__ huge, int hl) {
  const int s = blk_scene[npx_slot() / NPX_BLOCK];
  if (s < 0) return;
  const SceneC& S = scenes[s];
  const int q = npx_slot();
  if (q >= S.seg_off + S.n) return;
  const int shift = cstart[S.cls_base] - S.seg_off;
  const int c = cls_of[q];
  const int o = off_of[q];
  const int t = cstart[c] - shift + o;
  tperm[t] = q;
  tkey[t] = id[q];
  if (o == 0 && is_huge(t, cstart[c + 1] - shift, hl)) {
    const int h = atomicAdd(huge, 1);
    hid[c] = h < NPX_HUGE_MAX ? h : -1;
    if (h < NPX_HUGE_MAX) {
      huge[1 + 3 * h] = s;
      huge[2 + 3 * h] = t;
      huge[3 + 3 * h] = cstart[c + 1] - shift;
    }
  }
}


// Pass 3a: inside each 256-slot block holding part of a class larger than
// NPX_RANK_DIRECT, sorts the provisional slots by (class, original index).
// Slots are already grouped by class in ascending class order, so every
// class keeps its slot range and the part of it inside the block becomes
// sorted by index.  Bitonic sort in LDS.
__global__ void __launch_bounds__(NPX_BLOCK)
k_rank_sort(const SceneC* __restrict__ scenes, const int* __restrict__ blk_scene, const int* __restrict__ cstart,
            const int* __restrict__ cls_of, int* __restrict__ tperm, int* __restrict__ tkey,
            unsigned char* __restrict__ bigb, const int* __restrict__ hid, int hl) {
  __shared__ unsigned long long sk[NPX_BLOCK];
  __shared__ int sp[NPX_BLOCK];
  const int s = blk_scene[blockIdx.x];
  if (s < 0) return;  // uniform per block
  __shared__ int slv;
  const SceneC& S = scenes[s];
  const int i = (int)threadIdx.x;
  const int t = npx_slot();
  const bool active = t < S.seg_off + S.n;
  if (i == 0) slv = 0;
  int p = 0, c = 0, lv = 0;
  bool large = false;
  if (active) {
    p = tperm[t];
    c = cls_of[p];
    const int shift = cstart[S.cls_base] - S.seg_off;
    const int b = cstart[c] - shift, e = cstart[c + 1] - shift;
    large = e - b > NPX_RANK_DIRECT && !(is_huge(b, e, hl) && hid[c] >= 0);  // ranked in index space
    if (large) lv = merge_levels(b, e);
  }
  const bool big = __syncthreads_or(large) != 0;
  if (lv > 0) atomicMax(&slv, lv);
  __syncthreads();
  if (i == 0) bigb[blockIdx.x] = (unsigned char)slv;  // read by the merge launches
  if (!big) return;
  unsigned long long key = ~0ull;  // inactive slots (a suffix of the block) sort last
  if (active) key = ((unsigned long long)(unsigned)c << 32) | (unsigned)tkey[t];
  sk[i] = key;
  sp[i] = p;
  __syncthreads();
  for (int size = 2; size <= NPX_BLOCK; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      const int j = i ^ stride;
      if (j > i) {
        const unsigned long long a = sk[i], b = sk[j];
        if ((a > b) == ((i & size) == 0)) {
          sk[i] = b;
          sk[j] = a;
          const int tp = sp[i];
          sp[i] = sp[j];
          sp[j] = tp;
        }
      }
      __syncthreads();
    }
  }
  if (active) {
    tkey[t] = (int)(unsigned)sk[i];
    tperm[t] = sp[i];
  }
}

// Pass 3b: final slot = class start + rank of the particle's original index
// inside its class, so each bucket lists particles in ascending index like
// the reference's push_back order (neighbor_search.cpp:18-21).  Indices are
// unique within a scene.  Small classes (at most NPX_RANK_DIRECT) count
// directly.  A larger class lying inside one block is already sorted by
// k_rank_sort; one spanning several blocks is a sequence of sorted pieces
// (its parts of consecutive blocks) that the merge launches below combine.
// The class of each sorted slot goes to tcls: the merge launches' slot / key
// buffers hold valid entries only inside the ranges of the classes they
// merge.
#define NPX_RANK_STAGE 8

__global__ void __launch_bounds__(NPX_BLOCK)
k_rank(const SceneC* __restrict__ scenes, const int* __restrict__ blk_scene, const int* __restrict__ cstart,
       const int* __restrict__ cls_of, const int* __restrict__ tperm, const int* __restrict__ tkey,
       int* __restrict__ perm, int* __restrict__ tcls, const int* __restrict__ hid,
       unsigned char* __restrict__ hmark, int* __restrict__ hidp, int hl) {
  const int s = blk_scene[blockIdx.x];
  if (s < 0) return;  // uniform per block
  const SceneC& S = scenes[s];
  const int t = npx_slot();
  if (t >= S.seg_off + S.n) return;
  const int shift = cstart[S.cls_base] - S.seg_off;
  const int p = tperm[t];
  const int c = cls_of[p];
  tcls[t] = c;  // the class of every sorted slot, for the merge launches
  const int b = cstart[c] - shift, e = cstart[c + 1] - shift;
  // every original index of the scene gets its mark: the huge class's index
  // h, or none
  const int h = is_huge(b, e, hl) ? hid[c] : -1;
  const int ix = S.seg_off + tkey[t];
  hmark[ix] = h >= 0 ? (unsigned char)h : (unsigned char)NPX_HUGE_NONE;
  if (h >= 0) {
    hidp[ix] = p;  // placed by k_huge_place
  } else if (e - b <= NPX_RANK_DIRECT) {
    const int my = tkey[t];
    int r = 0;
    for (int u = b; u < e; ++u) r += (tkey[u] < my) ? 1 : 0;
    perm[b + r] = p;
  } else if (b / NPX_BLOCK == (e - 1) / NPX_BLOCK) {
    perm[t] = p;  // sorted by k_rank_sort
  }
}

// Merge launches of the large classes spanning several blocks (a segmented
// merge sort).  The class's pieces are its parts of consecutive 256-slot
// blocks, each sorted by k_rank_sort.  Before the launch at `level`, runs of
// 2^level consecutive pieces are sorted; the launch merges each group of
// NPX_MERGE_FAN consecutive runs (two levels of a pairwise merge sort in one
// launch): a slot moves to its group's start + its offset in its run + the
// number of smaller keys in the group's other runs (one binary search per
// run, the searches stepped together so their loads overlap; keys are unique
// within a class).  A class spanning G blocks is sorted after
// ceil(log4 G) launches.
#define NPX_MERGE_LOG 2
#define NPX_MERGE_FAN (1 << NPX_MERGE_LOG)

__device__ __forceinline__ int merge_dst(int b, int e, int t, int level, int my, const int* __restrict__ kin) {
  constexpr int F = NPX_MERGE_FAN - 1;
  const int B0 = b / NPX_BLOCK;
  auto run_lo = [&](int r) { return max(b, (B0 + (r << level)) * NPX_BLOCK); };
  auto run_hi = [&](int r) { return min(e, (B0 + ((r + 1) << level)) * NPX_BLOCK); };
  const int r = (t / NPX_BLOCK - B0) >> level;
  const int g = r & ~(NPX_MERGE_FAN - 1);
  int lo[F], hi[F], q0[F];
#pragma unroll
  for (int u = 0; u < F; ++u) {
    const int q = g + u + ((g + u >= r) ? 1 : 0);  // the group's runs other than r
    lo[u] = run_lo(q);
    hi[u] = run_hi(q);  // below lo when the run starts past the class
    q0[u] = lo[u];
  }
  for (;;) {  // lower bounds of my key
    bool any = false;
    int m[F], kv[F];
#pragma unroll
    for (int u = 0; u < F; ++u) {
      m[u] = (lo[u] + hi[u]) >> 1;
      kv[u] = (lo[u] < hi[u]) ? kin[m[u]] : my;
    }
#pragma unroll
    for (int u = 0; u < F; ++u) {
      if (lo[u] < hi[u]) {
        any = true;
        if (kv[u] < my)
          lo[u] = m[u] + 1;
        else
          hi[u] = m[u];
      }
    }
    if (!any) break;
  }
  int rank = 0;
#pragma unroll
  for (int u = 0; u < F; ++u) rank += (q0[u] < e) ? lo[u] - q0[u] : 0;
  return run_lo(g) + (t - run_lo(r)) + rank;
}

// One merge launch over the fluid sort's slots.  The launch that leaves one
// run writes the final permutation; earlier ones write the other key / slot
// buffers.  Blocks whose classes need no merge at this level (bigb, from
// k_rank_sort) exit at once, so idle launches cost one byte per block.
__global__ void __launch_bounds__(NPX_BLOCK)
k_rank_merge(const SceneC* __restrict__ scenes, const int* __restrict__ blk_scene, const int* __restrict__ cstart,
             const int* __restrict__ tcls, const unsigned char* __restrict__ bigb, int level,
             const int* __restrict__ pin, const int* __restrict__ kin, int* __restrict__ pout,
             int* __restrict__ kout, int* __restrict__ perm, const int* __restrict__ hid, int hl) {
  const int s = blk_scene[blockIdx.x];
  if (s < 0 || bigb[blockIdx.x] <= level) return;  // uniform per block
  const SceneC& S = scenes[s];
  const int t = npx_slot();
  if (t >= S.seg_off + S.n) return;
  const int shift = cstart[S.cls_base] - S.seg_off;
  const int c = tcls[t];
  const int b = cstart[c] - shift, e = cstart[c + 1] - shift;
  if (e - b <= NPX_RANK_DIRECT) return;
  const int npieces = (e - 1) / NPX_BLOCK - b / NPX_BLOCK + 1;
  if (npieces <= (1 << level)) return;  // merged by an earlier launch (or a single piece)
  if (is_huge(b, e, hl) && hid[c] >= 0) return;  // ranked in index space
  const int p = pin[t];
  const int my = kin[t];
  const int dst = merge_dst(b, e, t, level, my, kin);
  if (npieces <= (NPX_MERGE_FAN << level)) {
    perm[dst] = p;  // the last launch of this class
  } else {
    pout[dst] = p;
    kout[dst] = my;
  }
}

// Index-space rank of the huge classes.  Original indices are unique in
// [0, n) within a scene, and a class lists its particles in ascending index,
// so a huge class's particle goes to the class's first slot + the number of
// its members with a smaller index.  k_rank marked each index of a scene
// with its huge class (hmark[seg_off + index], NPX_HUGE_NONE otherwise) and
// the particle's unsorted slot (hidp).  Tiles are the 256 indices of one
// slot block.
//
// Histogram of one tile (block-wide): members of each huge class among the
// tile's indices, into hcnt[h * ntiles + tile] for every h < nh.
__device__ __forceinline__ void huge_hist_tile(int nh, bool valid, int ix, const unsigned char* __restrict__ hmark,
                                               int ntiles, int* __restrict__ hcnt) {
  __shared__ int cnt[NPX_HUGE_MAX];
  const int i = (int)threadIdx.x;
  if (i < NPX_HUGE_MAX) cnt[i] = 0;
  __syncthreads();
  if (valid) {
    const int m = hmark[ix];
    if (m < nh) atomicAdd(&cnt[m], 1);
  }
  __syncthreads();
  if (i < nh) hcnt[(size_t)i * ntiles + blockIdx.x] = cnt[i];
}

// k_huge_hist: members of each huge class per tile (every tile of a scene
// writes every row; a row's scan reads only the tiles of its own scene).
__global__ void __launch_bounds__(NPX_BLOCK)
k_huge_hist(const SceneC* __restrict__ scenes, const int* __restrict__ blk_scene, const int* __restrict__ huge,
            const unsigned char* __restrict__ hmark, int ntiles, int* __restrict__ hcnt) {
  const int nh = min(huge[0], NPX_HUGE_MAX);
  const int s = blk_scene[blockIdx.x];
  if (nh == 0 || s < 0) return;  // uniform per block
  const SceneC& S = scenes[s];
  const int ix = npx_slot();
  huge_hist_tile(nh, ix < S.seg_off + S.n, ix, hmark, ntiles, hcnt);
}

// k_huge_scan: one block per huge class, an exclusive scan of its row over
// the tiles of its scene (all ntiles tiles for an auxiliary sort's class,
// whose scene entry is -1).
__global__ void __launch_bounds__(1024)
k_huge_scan(const SceneC* __restrict__ scenes, const int* __restrict__ huge, int ntiles, int* __restrict__ hcnt) {
  __shared__ int wsum[16];
  const int h = (int)blockIdx.x;
  if (h >= min(huge[0], NPX_HUGE_MAX)) return;  // uniform per block
  const int s = huge[1 + 3 * h];
  int t0 = 0, t1 = ntiles;
  if (s >= 0) {
    t0 = scenes[s].seg_off / NPX_BLOCK;
    t1 = t0 + (scenes[s].n + NPX_BLOCK - 1) / NPX_BLOCK;
  }
  int* row = hcnt + (size_t)h * ntiles;
  const int lane = (int)(threadIdx.x & 63), wv = (int)(threadIdx.x >> 6);
  int carry = 0;
  for (int base = t0; base < t1; base += 1024) {
    const int k = base + (int)threadIdx.x;
    const int v = k < t1 ? row[k] : 0;
    int incl = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int u = __shfl_up(incl, o, 64);
      if (lane >= o) incl += u;
    }
    if (lane == 63) wsum[wv] = incl;
    __syncthreads();
    int pre = carry, tot = 0;
    for (int w = 0; w < 16; ++w) {
      const int x = wsum[w];
      pre += (w < wv) ? x : 0;
      tot += x;
    }
    if (k < t1) row[k] = pre + incl - v;
    carry += tot;
    __syncthreads();  // wsum is rewritten by the next chunk
  }
}

// Final slot of a huge class member (block-wide, one index per thread) =
// the class's first slot + members in earlier tiles + members at lower
// indices of this tile (earlier waves through LDS, earlier lanes by ballot);
// -1 for an index of no huge class.
__device__ __forceinline__ int huge_place_tile(int nh, bool valid, int ix, const int* __restrict__ huge,
                                               const unsigned char* __restrict__ hmark, int ntiles,
                                               const int* __restrict__ hcnt) {
  __shared__ int wc[NPX_BLOCK / 64][NPX_HUGE_MAX];
  const int i = (int)threadIdx.x;
  for (int k = i; k < (NPX_BLOCK / 64) * NPX_HUGE_MAX; k += NPX_BLOCK) (&wc[0][0])[k] = 0;
  __syncthreads();
  int m = valid ? (int)hmark[ix] : NPX_HUGE_NONE;
  if (m >= nh) m = NPX_HUGE_NONE;
  const int lane = i & 63, wv = i >> 6;
  const unsigned long long below = (1ull << lane) - 1ull;
  unsigned long long rest = __ballot(m != NPX_HUGE_NONE);
  int r = 0;
  while (rest) {  // one round per distinct huge class among the wave's indices
    const int v = __shfl(m, (int)__builtin_ctzll(rest), 64);
    const unsigned long long same = __ballot(m == v);
    if (m == v) r = __popcll(same & below);
    if (lane == 0) wc[wv][v] = __popcll(same);
    rest &= ~same;
  }
  __syncthreads();
  if (m == NPX_HUGE_NONE) return -1;
  for (int w = 0; w < wv; ++w) r += wc[w][m];
  const int dst = huge[2 + 3 * m] + hcnt[(size_t)m * ntiles + blockIdx.x] + r;
  return dst < huge[3 + 3 * m] ? dst : -1;
}

__global__ void __launch_bounds__(NPX_BLOCK)
k_huge_place(const SceneC* __restrict__ scenes, const int* __restrict__ blk_scene, const int* __restrict__ huge,
             const unsigned char* __restrict__ hmark, const int* __restrict__ hidp, int ntiles,
             const int* __restrict__ hcnt, int* __restrict__ perm) {
  const int nh = min(huge[0], NPX_HUGE_MAX);
  const int s = blk_scene[blockIdx.x];
  if (nh == 0 || s < 0) return;  // uniform per block
  const SceneC& S = scenes[s];
  const int ix = npx_slot();
  // marks are this step's: k_rank writes every index of the scene
  const int dst = huge_place_tile(nh, ix < S.seg_off + S.n, ix, huge, hmark, ntiles, hcnt);
  if (dst >= 0) perm[dst] = hidp[ix];
}

// Pass 4: reorder the state into the new slot order and form x*; also marks
// runs of identical state.  Every quantity a step computes for a particle is
// a function of its committed x, its v' = v + g dt and its cell's candidate
// order, so particles with bitwise equal (x, v') in one cell (equal x puts
// them in one cell, and the sort puts them in one class) get bitwise equal
// values in every pass, and each contributes the same term to a neighbour's
// sums.  A slot starts a run (rstart[q] = 1; the run's head) unless it holds
// the same (x, v') as the slot before it in the same scene; bcount[block]
// counts the block's run starts, the bit of a class in runb and
// runs_s[scene] are set when the class (the scene) holds a run longer than
// one.  k_pieces turns these into run lists.  Only heads compute in the
// neighbour passes: every read of a candidate's data goes to its run's head
// (the inline sweeps step run by run and add a run's term once, k times over,
// with npx_rep_add; list entries store heads), and k_run_fill copies the
// heads' final state to the other members.  Scenes with no solver iteration
// carry density from the previous step into the velocity update, so they get
// no runs, and runs_on = 0 (the FMA kernels) turns runs off.
__device__ __forceinline__ bool same_state(const float4& a, const float4& b, const float4& x, const float4& w) {
  return __float_as_uint(a.x) == __float_as_uint(x.x) && __float_as_uint(a.y) == __float_as_uint(x.y) &&
         __float_as_uint(a.z) == __float_as_uint(x.z) && __float_as_uint(b.x) == __float_as_uint(w.x) &&
         __float_as_uint(b.y) == __float_as_uint(w.y) && __float_as_uint(b.z) == __float_as_uint(w.z);
}

__global__ void __launch_bounds__(NPX_BLOCK)
k_gather(const SceneC* __restrict__ scenes, const int* __restrict__ blk_scene, const int* __restrict__ perm,
         const float4* __restrict__ X, const float4* __restrict__ V, const int* __restrict__ id,
         const int* __restrict__ cell_of, const int* __restrict__ clsmap, float4* __restrict__ X2,
         float4* __restrict__ V2, float4* __restrict__ XS, int* __restrict__ id2, int* __restrict__ cell2,
         int runs_on, unsigned char* __restrict__ rstart, unsigned* __restrict__ runb, int* __restrict__ runs_s,
         int* __restrict__ bcount) {
  __shared__ float4 sx[NPX_BLOCK], sv[NPX_BLOCK];
  __shared__ unsigned long long smask[NPX_BLOCK / 64];
  __shared__ int scount;
  const int t = (int)threadIdx.x;
  if (blockIdx.x == 0 && t == 0) bcount[gridDim.x] = 0;  // the scan's last entry
  const int s = blk_scene[blockIdx.x];
  if (s < 0) {  // uniform per block
    if (t == 0) bcount[blockIdx.x] = 0;
    return;
  }
  const SceneC& S = scenes[s];
  const int q = npx_slot();
  const bool active = q < S.seg_off + S.n;
  float4 x = make_float4(0.0f, 0.0f, 0.0f, 0.0f), w = x;
  int c = 0;
  if (t == 0) scount = 0;
  if (active) {
    const int p = perm[q];
    x = X[p];
    const float4 v = V[p];
    const float vx = v.x + S.gdt0, vy = v.y + S.gdt1, vz = v.z + S.gdt2;
    w = make_float4(vx, vy, vz, 0.0f);
    X2[q] = x;
    V2[q] = w;
    XS[q] = make_float4(x.x + vx * S.dt, x.y + vy * S.dt, x.z + vz * S.dt, 0.0f);
    id2[q] = id[p];
    c = cell_of[p];
    cell2[q] = c;
  }
  sx[t] = x;
  sv[t] = w;
  __syncthreads();
  bool start = active;  // inactive slots (a suffix of the block) are not counted
  if (active && S.iters > 0 && runs_on && q > S.seg_off) {
    if (t > 0) {  // active slots form a prefix of the block, so t - 1 is active
      start = !same_state(sx[t - 1], sv[t - 1], x, w);
    } else {  // the last slot of the previous block (same scene, active)
      const int p = perm[q - 1];
      const float4 v = V[p];
      start = !same_state(X[p], make_float4(v.x + S.gdt0, v.y + S.gdt1, v.z + S.gdt2, 0.0f), x, w);
    }
  }
  const unsigned long long m = __ballot(start);
  if ((t & 63) == 0) {
    smask[t >> 6] = m;
    atomicAdd(&scount, (int)__popcll(m));
  }
  __syncthreads();
  if (active) {
    rstart[q] = start ? 1 : 0;
    // a run longer than one: flagged by its second slot (or by the block's
    // first slot when the run comes from the previous block)
    const int lane = t & 63;
    const bool prev_start = (t == 0) || ((lane > 0) ? ((m >> (lane - 1)) & 1ull) : (smask[(t >> 6) - 1] >> 63));
    if (!start && prev_start) {
      const unsigned cl = (unsigned)clsmap[S.cls_base + c];
      atomicOr(runb + (cl >> 5), 1u << (cl & 31u));
      runs_s[s] = 1;
    }
  }
  if (t == 0) bcount[blockIdx.x] = scount;
}

// Run lists, after the exclusive scan of bcount into boff: piece i (run i in
// slot order) starts at slot ps[i], and pi[q] is the piece of slot q's run.
// The runs of a slot range [b, e) are pieces pi[b] .. pi[e - 1] (b is always
// a run start: a class start differs from the slot before it in cell), the
// last ending at e.
//
// Block 0 also sets the step's XCD order (order[0]): the run heads, the
// particles that compute in the neighbour passes, per eighth of the slot
// range; when the fullest eighth holds more than 5/4 of the mean, one
// contiguous share per XCD would leave most XCDs idle (a collapsed scene
// keeps its heads in a few places), and the passes deal chunks to the XCDs
// in turn instead (1); else contiguous shares (0).
__global__ void __launch_bounds__(NPX_BLOCK)
k_pieces(const SceneC* __restrict__ scenes, const int* __restrict__ blk_scene,
         const unsigned char* __restrict__ rstart, const int* __restrict__ boff, int* __restrict__ pi,
         int* __restrict__ ps, int* __restrict__ order) {
  __shared__ unsigned long long smask[NPX_BLOCK / 64];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    const int nb = (int)gridDim.x;
    int mx = 0;
    for (int x = 0; x < 8; ++x) {
      const int a = (int)((long long)nb * x / 8), b = (int)((long long)nb * (x + 1) / 8);
      mx = max(mx, boff[b] - boff[a]);
    }
    order[0] = (32LL * mx > 5LL * boff[nb]) ? 1 : 0;
  }
  const int s = blk_scene[blockIdx.x];
  if (s < 0) return;  // uniform per block
  const SceneC& S = scenes[s];
  const int t = (int)threadIdx.x;
  const int q = npx_slot();
  const bool active = q < S.seg_off + S.n;
  const bool st = active && rstart[q];
  const unsigned long long m = __ballot(st);
  if ((t & 63) == 0) smask[t >> 6] = m;
  __syncthreads();
  if (!active) return;
  const int lane = t & 63;
  int before = (int)__popcll((lane == 63) ? m : (m & ((2ull << lane) - 1ull)));  // starts at or before q
  for (int u = 0; u < (t >> 6); ++u) before += (int)__popcll(smask[u]);
  const int i = boff[blockIdx.x] + before - 1;  // members before the block's first start: the previous piece
  pi[q] = i;
  if (st) ps[i] = q;
}

// Carries density and lambda into the new slot order (only needed when a
// scene runs zero solver iterations and keeps the previous values).
__global__ void __launch_bounds__(NPX_BLOCK)
k_gather_scalar2(const SceneC* __restrict__ scenes, const int* __restrict__ blk_scene, const int* __restrict__ perm,
                 const float* __restrict__ A, const float* __restrict__ B, float* __restrict__ A2,
                 float* __restrict__ B2) {
  const int s = blk_scene[npx_slot() / NPX_BLOCK];
  if (s < 0) return;
  const SceneC& S = scenes[s];
  const int q = npx_slot();
  if (q >= S.seg_off + S.n) return;
  const int p = perm[q];
  A2[q] = A[p];
  B2[q] = B[p];
}

// ------------------------------------------------------ PBF passes
#define NPX_SCENE_PROLOGUE()                                 \
  const int s = blk_scene[npx_slot() / NPX_BLOCK];           \
  if (s < 0) return;                                         \
  const SceneC& S = scenes[s];                               \
  const int q = npx_slot();                                  \
  if (q >= S.seg_off + S.n) return;                          \
  const int shift = cstart[S.cls_base] - S.seg_off;          \
  const int wxy = S.wx * S.wy;                               \
  const int* __restrict__ cm = clsmap + S.cls_base;          \
  const int l = cell[q];

// Block order of the neighbour passes.  The hardware hands workgroups to the
// 8 XCDs round robin (workgroup i to XCD i mod 8), and each XCD has its own
// L2.  With `remap` (the XCD order) launch block i processes logical block
//     x * (G / 8) + min(x, G mod 8) + i / 8,   x = i mod 8,
// a bijection of [0, G) that gives XCD x one contiguous run of slots, so the
// candidate rows its blocks share stay in its L2.  Results do not depend on
// the order: no pass reads what another block of the same launch writes.
struct BlockPlan {
  int remap;  // 1: XCD order (contiguous shares or chunks, per *order); 0: hardware order; C >= 2: chunks of C
  int nd;     // leading blocks that sweep the dense list (a multiple of 8)
  int nvb;    // particle blocks of the pass
  const int* order;  // this step's XCD order (k_pieces): 0 contiguous shares, 1 chunks
};

// Slots per chunk of the chunked XCD order.
#define NPX_XCD_CHUNK_SLOTS 8192

// The particle block of a neighbour-pass launch block (after the nd dense
// blocks, which keep XCD i mod 8).
__device__ __forceinline__ int npx_lblock(const BlockPlan& bp) {
  const int i = (int)blockIdx.x - bp.nd;
  if (!bp.remap) return i;
  const int g = (int)gridDim.x - bp.nd, x = i & 7;
  int C = bp.remap;
  if (C == 1) {
    if (*bp.order == 0) return x * (g >> 3) + min(x, g & 7) + (i >> 3);
    C = max(NPX_XCD_CHUNK_SLOTS / (int)blockDim.x, 1);
  }
  // chunks of C consecutive blocks dealt to the XCDs in turn (the tail past
  // the last whole round in hardware order)
  const int full = g - g % (8 * C);
  if (i >= full) return i;
  const int j = i >> 3;
  return ((j / C) * 8 + x) * C + (j % C);
}

// Per-scene list statistics (Skin::lst), under list policy 1 (the default,
// 0, builds no skin lists: with the two-phase inline sweep they measured
// slower on every workload tried): the testing passes count the lanes that
// test and those whose list bound (cd + C <= skin_allow) held, whether or not
// a list was built; the next step builds skin lists only when at least a
// quarter held (or nothing was counted yet).  Where the candidates move
// farther than the skin per iteration, almost every particle sweeps inline
// anyway, and building the lists would only cost their writes.  Both ways
// give the same floats: a particle without a list sweeps inline.
// The counts are a sample (every eighth wave of a testing pass) spread over
// NPX_LST_COPIES counter pairs, 64 bytes apart, so the atomics of a pass do
// not queue on one cache line.  Per scene: NPX_LST_COPIES x 16 ints of counter
// pairs (tested, held), then the build flag.
#define NPX_LST_COPIES 32
#define NPX_LST_BUILD (16 * NPX_LST_COPIES)
#define NPX_LST_STRIDE (NPX_LST_BUILD + 16)

// Empties the step's dense list, which the build pass fills after it, and
// turns the previous step's list statistics into this step's build flags.
__global__ void k_dense_reset(int* __restrict__ dn, int* __restrict__ lst, int nscenes, int policy) {
  if (threadIdx.x == 0) *dn = 0;
  for (int s = (int)threadIdx.x; s < nscenes; s += (int)blockDim.x) {
    int* st = lst + (size_t)NPX_LST_STRIDE * s;
    long long tested = 0, held = 0;
    for (int k = 0; k < NPX_LST_COPIES; ++k) {
      tested += st[16 * k];
      held += st[16 * k + 1];
      st[16 * k] = 0;
      st[16 * k + 1] = 0;
    }
    st[NPX_LST_BUILD] = (policy == 1 && (tested == 0 || 4 * held >= tested)) ? 1 : 0;
  }
}

// Neighbour-pass prologue.  `active` lanes own a particle; the others stay
// alive to the end (the delta pass reduces over whole waves).
// Runs f(lb) for the particle blocks of a neighbour-pass launch (the launch
// blocks after the nd dense ones): one launch block per particle block, in
// XCD order (npx_lblock).  The persistent instantiations (PERSIST) are launched
// with fewer blocks (a multiple of 8, set by npx_set_persistent), and each
// loops over particle blocks; in XCD order the blocks of XCD x take the
// particle blocks of the x-th contiguous chunk in turn, so each XCD still
// sweeps one run of particles.  (A separate instantiation: the loop would
// otherwise keep values live across particle blocks and make the one-shot
// kernels spill.)
template <bool PERSIST, class F>
__device__ __forceinline__ void npx_for_blocks(const BlockPlan& bp, F&& f) {
  if constexpr (!PERSIST) {
    f(npx_lblock(bp));
    return;
  }
  const int i = (int)blockIdx.x - bp.nd, g = (int)gridDim.x - bp.nd, V = bp.nvb;
  if (bp.remap != 1) {  // (chunked orders: hardware order)
    for (int lb = i; lb < V; lb += g) f(lb);
    return;
  }
  const int x = i & 7, per = g >> 3;
  const int cs = (V >> 3) + ((x < (V & 7)) ? 1 : 0), c0 = x * (V >> 3) + min(x, V & 7);
  for (int k = i >> 3; k < cs; k += per) f(c0 + k);
}

// Neighbour-pass prologue for particle block lb.
#define NPX_NBR_PROLOGUE(lb)                                 \
  const int q = (lb) * (int)blockDim.x + (int)threadIdx.x;   \
  const int s = blk_scene[q / NPX_BLOCK];                    \
  if (s < 0) return; /* uniform per block */                 \
  const SceneC& S = scenes[s];                               \
  const bool active = q < S.seg_off + S.n;                   \
  const int shift = cstart[S.cls_base] - S.seg_off;          \
  const int wxy = S.wx * S.wy;                               \
  const int* __restrict__ cm = clsmap + S.cls_base;          \
  const int l = active ? cell[q] : 0;

// Skin lists.  The first density pass of a step sweeps every candidate (the
// 27-cell sweep, reference order) and keeps, per particle, the candidates
// within the skin radius h_s = (1 + f) h, in order.  Later passes of the step
// visit only that list.  A candidate j left out of i's list had |r| >= h_s at
// x*_0, and |r| can since have shrunk by at most c_i + c_j, c the summed
// per-iteration displacement of a particle.  The candidates of i are the
// particles of the alias classes of its 27 cells, so with C_i the largest c
// over those classes (kept per class with atomic maxima), i may use its list
// while c_i + C_i <= h_s - |cut| (skin_allow, with relative margins for
// rounding): then the list holds every in-range candidate.  Each testing pass (density,
// and the XSPH / curl pass at the final x*) decides per particle, tests the
// list entries with the cheap distance check, runs the SPH arithmetic over
// the in-range entries only and writes them as a compacted list; the pass
// that reuses the same positions (delta-p, vorticity force) follows that
// list, or sweeps inline when the testing pass found more in-range
// candidates than a list holds (in-range count -1).  Skipped entries are
// exact zero terms of the reference's sums, so the results are the
// reference's bit for bit, and every lane of a wave works on a real pair.  A
// particle with more than NPX_SKIN_CAP skin candidates, or one whose bound
// fails, sweeps its candidates inline (an inline sweep of a testing pass
// still records the in-range list when it fits), and a crowded one with an
// overflowing list goes to the per-wave dense kernels.  Both kinds of list
// are stored in chunks of four entries per slot (list_at).  Visitors fetch
// the data of G candidates before evaluating any of them, so G loads are in
// flight per wave.
#define NPX_ROWS 9
#define NPX_SKIN_CAP 64

struct Skin {
  int* __restrict__ idx;     // [slots / 64][NPX_SKIN_CAP][64] skin list
  int* __restrict__ cnt;     // [slots]; -1: sweep inline; NPX_CNT_DENSE: on the dense list
  int* __restrict__ ex;      // same layout: in-range list of the current testing pass
  int* __restrict__ ecnt;    // [slots]; -1: the testing pass swept inline
  float* __restrict__ cd;      // [slots] summed displacement since the lists were built
  unsigned* __restrict__ cmax; // [classes] float bits of the largest cd in each class
  const unsigned char* __restrict__ rstart;  // [slots] 1: the slot starts a run (a head)
  const int* __restrict__ pi;               // [slots] piece (run) of the slot
  const int* __restrict__ ps;               // [pieces] first slot of each run
  const unsigned* __restrict__ runb;       // [classes / 32] bit set: the class holds a run longer than one
  const int* __restrict__ runs_s;          // [scenes] 1: the scene holds a run longer than one
  int* __restrict__ dl;                    // [slots] dense list: leaders of batches of crowded heads
  int* __restrict__ dn;                    // [1] dense list length
  int dense;                               // 1: dense particles go to the per-wave kernels
  int* __restrict__ lst;                   // [scenes][NPX_LST_STRIDE]: sampled counts of testing lanes and of
                                           // lanes whose list bound held, whether this step builds lists
  int lstat;                               // 1: the testing passes sample those counts (list policy 1)
};

// Entry k of slot q's skin or in-range list.  Entries go in chunks of four
// per slot, and chunk c of the 64 slots of a wave forms one 1 KB row: a lane
// writes and reads four entries with one 16-byte access, and lanes at the
// same chunk share a row (the passes that build a list collect four entries
// per store; the passes that walk one, in lockstep, read one row per load).
__device__ __forceinline__ size_t list_at(int q, int k) {
  return ((((size_t)(q >> 6) * (NPX_SKIN_CAP / 4) + (k >> 2)) * 64 + (q & 63)) << 2) + (k & 3);
}

// Appends of one lane to a list in list_at layout, four entries per store;
// past NPX_SKIN_CAP entries nothing is stored (the list has overflowed).
struct ListWriter {
  int* __restrict__ list;
  int q;
  int n = 0;
  int4 buf = make_int4(0, 0, 0, 0);
  __device__ __forceinline__ ListWriter(int* __restrict__ list_, int q_) : list(list_), q(q_) {}
  __device__ __forceinline__ void push(int j) {
    const int c = n & 3;
    buf.x = (c == 0) ? j : buf.x;
    buf.y = (c == 1) ? j : buf.y;
    buf.z = (c == 2) ? j : buf.z;
    buf.w = (c == 3) ? j : buf.w;
    if (c == 3 && n < NPX_SKIN_CAP) *reinterpret_cast<int4*>(list + list_at(q, n - 3)) = buf;
    ++n;
  }
  // stores the last, partial chunk (entries past n in it are never read)
  __device__ __forceinline__ void flush() {
    if ((n & 3) && n <= NPX_SKIN_CAP) *reinterpret_cast<int4*>(list + list_at(q, n & ~3)) = buf;
  }
};

// Entries k .. k+3 of slot q's list (k a multiple of 4) with one 16-byte load;
// those at or past n (n > k) are replaced by entry k, so every index is one
// written this step.
__device__ __forceinline__ void list_read4(const int* __restrict__ list, int q, int k, int n, int* js) {
  const int4 w = *reinterpret_cast<const int4*>(list + list_at(q, k));
  js[0] = w.x;
  js[1] = (k + 1 < n) ? w.y : w.x;
  js[2] = (k + 2 < n) ? w.z : w.x;
  js[3] = (k + 3 < n) ? w.w : w.x;
}

// Largest summed displacement among the candidates of the particle in window
// cell l (the classes of its 27 cells).
__device__ __forceinline__ float candidate_disp(const Skin& K, int l, int wx, int wxy, const int* __restrict__ cm) {
  float mx = 0.0f;
#pragma unroll
  for (int r = 0; r < NPX_ROWS; ++r) {
    const int base = l + (r % 3 - 1) * wx + (r / 3 - 1) * wxy;
#pragma unroll
    for (int d = -1; d <= 1; ++d) mx = fmaxf(mx, __uint_as_float(K.cmax[cm[base + d]]));
  }
  return mx;
}

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  return v;
}

template <bool FAST>
__device__ __forceinline__ float dist2(float rx, float ry, float rz) {
  return FAST ? __builtin_fmaf(rx, rx, __builtin_fmaf(ry, ry, rz * rz)) : sqn3(rx, ry, rz);
}

// Pins loaded values in registers at this point, so the G gathers of a batch
// are all issued before the first pair is evaluated (the compiler would
// otherwise sink a load into the branch that uses it).  The asm statement is
// empty: it only consumes the registers.
__device__ __forceinline__ void npx_keep(const float4& a) {
  __asm__ volatile("" ::"v"(a.x), "v"(a.y), "v"(a.z), "v"(a.w));
}

__device__ __forceinline__ void keep(const float4& a) { npx_keep(a); }

// Whether any of the n (1 or 3) consecutive classes from c holds a run.
__device__ __forceinline__ bool cls_runs(const Skin& K, int c, int n) {
  const unsigned sh = (unsigned)c & 31u, mask = (1u << n) - 1u;
  const unsigned* w = K.runb + ((unsigned)c >> 5);
  unsigned bits = w[0] >> sh;
  if (sh + (unsigned)n > 32u) bits |= w[1] << (32u - sh);
  return (bits & mask) != 0u;
}

struct Rows {
  unsigned long long m[NPX_ROWS];  // skin: r^2 < skin_cut2
  unsigned long long e[NPX_ROWS];  // in range: r^2 < cut
  int b[NPX_ROWS];
  unsigned rf;                     // bit r: row r's classes hold runs
};

// Sweeps the nine candidate rows of a particle at xi; per row, bit t of m
// (of e) is set when candidate b + t is inside the skin (in range).  False
// when a row is not one range of at most 64 slots.
template <bool FAST>
__device__ __forceinline__ bool sweep_rows(Rows& w, const Skin& K, bool rs, float4 xi, float cut, float skin_cut,
                                           const float4* __restrict__ XS, int l, int wx, int wxy,
                                           const int* __restrict__ cm, const int* __restrict__ cstart, int shift) {
  bool ok = true;
  w.rf = 0u;
#pragma unroll
  for (int r = 0; r < NPX_ROWS; ++r) {
    const int base = l + (r % 3 - 1) * wx + (r / 3 - 1) * wxy;
    const int c0 = cm[base - 1];
    const int c1 = cm[base];
    const int c2 = cm[base + 1];
    const int b = cstart[c0] - shift;
    const int e = cstart[c0 + 3] - shift;
    ok = ok && (c1 == c0 + 1) && (c2 == c0 + 2) && (e - b <= 64);
    unsigned long long mk = 0, ek = 0;
    if (ok) {
      for (int j = b; j < e; j += 4) {
        // four slots from one base address, up to three past e - 1 (the x*
        // buffers carry NPX_SLOT_PAD elements of padding), masked
        const float4* __restrict__ p = XS + j;
        float4 a[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) a[u] = p[u];
        unsigned sb = 0, eb = 0;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const float d = dist2<FAST>(xi.x - a[u].x, xi.y - a[u].y, xi.z - a[u].z);
          const bool v = j + u < e;
          sb |= (unsigned)(v & (d < skin_cut)) << u;
          eb |= (unsigned)(v & (d < cut)) << u;
        }
        mk |= (unsigned long long)sb << (j - b);  // j - b <= 60
        ek |= (unsigned long long)eb << (j - b);
      }
    }
    w.m[r] = mk;
    w.e[r] = ek & mk;
    w.b[r] = b;
    if (rs && ok && cls_runs(K, c0, 3)) w.rf |= 1u << r;
  }
  return ok;
}

// Writes the skin entries of the row masks to the list of slot q, in order
// (consumes w); returns the entry count (> NPX_SKIN_CAP: overflow) and the
// mask of in-range entries.
__device__ __forceinline__ int store_skin(Rows& w, const Skin& K, int q, unsigned long long* rel_out) {
  unsigned long long mk = w.m[0], ek = w.e[0], rel = 0;
  int b = w.b[0];
  unsigned rf = w.rf;  // bit 0: the current row holds runs (its entries map to run heads)
  int left = NPX_ROWS - 1;
  ListWriter out(K.idx, q);
  for (;;) {
    while (mk == 0 && left > 0) {
      mk = w.m[1];
      ek = w.e[1];
      b = w.b[1];
      rf >>= 1;
#pragma unroll
      for (int r = 1; r < NPX_ROWS - 1; ++r) {
        w.m[r] = w.m[r + 1];
        w.e[r] = w.e[r + 1];
        w.b[r] = w.b[r + 1];
      }
      --left;
    }
    if (mk == 0) break;
    const int t = (int)__builtin_ctzll(mk);
    mk &= mk - 1;
    if (out.n < NPX_SKIN_CAP) rel |= ((ek >> t) & 1ull) << out.n;
    out.push((rf & 1u) ? K.ps[K.pi[b + t]] : b + t);  // the run's head holds the data
  }
  out.flush();
  *rel_out = rel;
  return out.n;
}

// Builds the skin list of slot q by sweeping every candidate in order (any
// row structure, aliased cells included) and appending those inside the
// skin; returns the entry count (> NPX_SKIN_CAP: overflow) and the mask of
// in-range entries.  Used for the particles whose rows are not single slot
// ranges of at most 64 candidates.  With RUNS, classes holding runs of
// identical state are stepped run by run (one distance test per run).
template <bool FAST, bool RUNS>
__device__ __forceinline__ int sweep_direct(const Skin& K, bool rs, int q, float4 xi, float cut, float skin_cut,
                                            const float4* __restrict__ XS, int l, int wx, int wxy,
                                            const int* __restrict__ cm, const int* __restrict__ cstart, int shift,
                                            unsigned long long* rel_out) {
  int k = 0;
  unsigned long long rel = 0;
  // visits the candidates in the reference's order (for_each_candidate) and
  // stops once the list has overflowed
#pragma unroll 1
  for (int r = 0; r < NPX_ROWS && k <= NPX_SKIN_CAP; ++r) {
    const int base = l + (r % 3 - 1) * wx + (r / 3 - 1) * wxy;
    const int c0 = cm[base - 1], c1 = cm[base], c2 = cm[base + 1];
    const bool one = (c1 == c0 + 1 && c2 == c0 + 2);
#pragma unroll 1
    for (int m = 0; m < (one ? 1 : 3) && k <= NPX_SKIN_CAP; ++m) {
      const int c = (m == 0) ? c0 : (m == 1) ? c1 : c2;
      const int b = cstart[c] - shift, e = cstart[one ? c0 + 3 : c + 1] - shift;
      if (RUNS && rs && b < e && (one ? cls_runs(K, c0, 3) : cls_runs(K, c, 1))) {
        const int iend = K.pi[e - 1] + 1;
        for (int i = K.pi[b]; i < iend && k <= NPX_SKIN_CAP; i += 2) {  // two runs per fetch
          const int j0 = K.ps[i];
          const int j1 = (i + 1 < iend) ? K.ps[i + 1] : e;
          const int j2 = (i + 2 < iend) ? K.ps[i + 2] : e;
          const float4 a0 = XS[j0], a1 = XS[(i + 1 < iend) ? j1 : j0];
          const float d0 = dist2<FAST>(xi.x - a0.x, xi.y - a0.y, xi.z - a0.z);
          const float d1 = dist2<FAST>(xi.x - a1.x, xi.y - a1.y, xi.z - a1.z);
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            const int j = h ? j1 : j0, kk = (h ? j2 : j1) - j;
            const float d = h ? d1 : d0;
            if ((h == 0 || i + 1 < iend) && k <= NPX_SKIN_CAP && d < skin_cut) {
              if (k + kk > NPX_SKIN_CAP) {
                k = NPX_SKIN_CAP + 1;  // overflow
              } else {
                for (int u = 0; u < kk; ++u, ++k) {
                  K.idx[list_at(q, k)] = j;  // the run's head
                  rel |= (unsigned long long)(d < cut) << k;
                }
              }
            }
          }
        }
      } else {
        for (int j = b; j < e && k <= NPX_SKIN_CAP; ++j) {
          const float4 a = XS[j];
          const float d = dist2<FAST>(xi.x - a.x, xi.y - a.y, xi.z - a.z);
          if (d < skin_cut) {
            if (k < NPX_SKIN_CAP) {
              K.idx[list_at(q, k)] = j;  // no runs here: every slot is a head
              rel |= (unsigned long long)(d < cut) << k;
            }
            ++k;
          }
        }
      }
    }
  }
  *rel_out = rel;
  return k;
}

// Mask of the skin-list entries of slot q (n of them) with r^2 < cut at xi.
template <bool FAST>
__device__ __forceinline__ unsigned long long test_list(const Skin& K, int q, int n, float4 xi, float cut,
                                                        const float4* __restrict__ XS) {
  unsigned long long rel = 0;
  for (int k = 0; k < n; k += 4) {
    int js[4];
    list_read4(K.idx, q, k, n, js);
    float4 a[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) a[u] = XS[js[u]];
    unsigned bits = 0;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const float d = dist2<FAST>(xi.x - a[u].x, xi.y - a[u].y, xi.z - a[u].z);
      bits |= (unsigned)((d < cut) & (k + u < n)) << u;
    }
    rel |= (unsigned long long)bits << k;  // k <= 60
  }
  return rel;
}

// f(j, ld(j)) over the skin-list entries of slot q whose bit is set in rel,
// in list order; writes them as the compacted in-range list and returns its
// length.
template <int G, class L, class F>
__device__ __forceinline__ int visit_rel(const Skin& K, int q, unsigned long long rel, L&& ld, F&& f) {
  int e = 0;
  for (;;) {
    int ks[G];
#pragma unroll
    for (int u = 0; u < G; ++u) {
      ks[u] = rel ? (int)__builtin_ctzll(rel) : -1;
      rel &= rel - 1;
    }
    if (ks[0] < 0) break;
    int js[G];
#pragma unroll
    for (int u = 0; u < G; ++u) js[u] = K.idx[list_at(q, ks[u] < 0 ? ks[0] : ks[u])];
    decltype(ld(0)) v[G];
#pragma unroll
    for (int u = 0; u < G; ++u) v[u] = ld(js[u]);
#pragma unroll
    for (int u = 0; u < G; ++u) keep(v[u]);
    // every group but the last appends G entries, so with G = 4 a group
    // starts at a chunk boundary and its entries go out as one chunk (entries
    // of a partial last group past the list's end are never read)
    if constexpr (G == 4) *reinterpret_cast<int4*>(K.ex + list_at(q, e)) = make_int4(js[0], js[1], js[2], js[3]);
#pragma unroll
    for (int u = 0; u < G; ++u) {
      if (ks[u] >= 0) {
        if constexpr (G != 4) K.ex[list_at(q, e)] = js[u];
        ++e;
        f(js[u], v[u], 1);
      }
    }
  }
  return e;
}

// f(j, ld(j)) over the n entries of slot q's in-range list (list_at layout).
// With G = 4 each round reads one chunk with a 16-byte load (k is a multiple
// of 4); entries past n belong to the partial last chunk or to earlier steps
// and are replaced by entry k before any use.
template <int G, class L, class F>
__device__ __forceinline__ void visit_list(const int* __restrict__ list, int q, int n, L&& ld, F&& f) {
  for (int k = 0; k < n; k += G) {
    int js[G];
    if constexpr (G == 4) {
      list_read4(list, q, k, n, js);
    } else {
#pragma unroll
      for (int u = 0; u < G; ++u) js[u] = list[list_at(q, min(k + u, n - 1))];
    }
    decltype(ld(0)) v[G];
#pragma unroll
    for (int u = 0; u < G; ++u) v[u] = ld(js[u]);
#pragma unroll
    for (int u = 0; u < G; ++u) keep(v[u]);
#pragma unroll
    for (int u = 0; u < G; ++u)
      if (k + u < n) f(js[u], v[u], 1);
  }
}

template <int G, class L, class F>
__device__ __forceinline__ void visit_range(int b, int e, L&& ld, F&& f) {
  int j = b;
  for (; j + G <= e; j += G) {
    decltype(ld(0)) v[G];
#pragma unroll
    for (int u = 0; u < G; ++u) v[u] = ld(j + u);
#pragma unroll
    for (int u = 0; u < G; ++u) keep(v[u]);
#pragma unroll
    for (int u = 0; u < G; ++u) f(j + u, v[u], 1);
  }
  for (; j < e; ++j) f(j, ld(j), 1);
}

// visit_range over a slot range holding runs of identical state: one call
// per run, f(j, ld(j), k) standing for the k equal terms of slots j..j+k-1,
// with the starts and data of G runs fetched before any is evaluated.
template <int G, class L, class F>
__device__ __forceinline__ void visit_runs(const Skin& K, int b, int e, L&& ld, F&& f) {
  if (b >= e) return;
  const int iend = K.pi[e - 1] + 1;
  for (int i = K.pi[b]; i < iend; i += G) {
    int js[G + 1];
#pragma unroll
    for (int u = 0; u <= G; ++u) js[u] = (i + u < iend) ? K.ps[i + u] : e;
    decltype(ld(0)) v[G];
#pragma unroll
    for (int u = 0; u < G; ++u) v[u] = ld((i + u < iend) ? js[u] : js[0]);
#pragma unroll
    for (int u = 0; u < G; ++u) keep(v[u]);
#pragma unroll
    for (int u = 0; u < G; ++u)
      if (i + u < iend) f(js[u], v[u], js[u + 1] - js[u]);
  }
}

// f(j, ld(j), k) over every candidate in the reference's order (the 27-cell
// sweep of for_each_candidate); f applies the cut itself.  With RUNS, classes
// flagged by k_gather are walked run by run (k = run length, RG runs fetched
// at a time), so a stack of
// m coincident particles costs one term per run instead of m.
template <int G, bool RUNS, int RG = 2, class L, class F>
__device__ __forceinline__ void visit_all(const Skin& K, bool rs, int l, int wx, int wxy, const int* __restrict__ cm,
                                          const int* __restrict__ cstart, int shift, L&& ld, F&& f) {
#pragma unroll 1
  for (int r = 0; r < NPX_ROWS; ++r) {
    const int base = l + (r % 3 - 1) * wx + (r / 3 - 1) * wxy;
    const int c0 = cm[base - 1];
    const int c1 = cm[base];
    const int c2 = cm[base + 1];
    if (c1 == c0 + 1 && c2 == c0 + 2) {
      const int b = cstart[c0] - shift, e = cstart[c0 + 3] - shift;
      if (RUNS && rs && cls_runs(K, c0, 3))
        visit_runs<RG>(K, b, e, ld, f);
      else
        visit_range<G>(b, e, ld, f);
    } else {
#pragma unroll 1
      for (int k = 0; k < 3; ++k) {
        const int c = (k == 0) ? c0 : (k == 1) ? c1 : c2;
        const int b = cstart[c] - shift, e = cstart[c + 1] - shift;
        if (RUNS && rs && cls_runs(K, c, 1))
          visit_runs<RG>(K, b, e, ld, f);
        else
          visit_range<G>(b, e, ld, f);
      }
    }
  }
}

// Compacted visit for the inline sweep: slots are tested 64 at a time (r^2 <
// cut at xi, the test every pass body applies first) and only the in-range
// ones are fetched again and evaluated, in order.  A wave whose lanes sweep
// dense cells then evaluates each lane's in-range candidates side by side
// instead of stepping every lane through every candidate.

// Mask of the slots c .. c + m - 1 (1 <= m <= 64) with r^2 < cut at xi, or
// a superset of it: every caller re-tests r^2 < cut exactly before a slot
// adds anything (the pair bodies return on r^2 >= cut, or on the spiky
// support bound sp_hi <= cut for the vorticity force; the in-range recording
// re-tests), so extra slots only cost a visit.  The exact test evaluates
// x*x + (y*y + z*z) with three roundings; here r^2 is one FMA chain (also
// within 3 ulp of the true value), compared against cut enlarged by 2^-20,
// which admits every slot the exact test admits.  Slots are read four at a
// time from one base address, up to three past c + m - 1 (the x* buffers
// carry NPX_SLOT_PAD elements of padding), and masked.
template <bool FAST>
__device__ __forceinline__ unsigned long long test_slots(int c, int m, float4 xi, float cut,
                                                         const float4* __restrict__ XS) {
  unsigned long long rel = 0;
  const float ct = FAST ? cut : cut * (1.0f + 0x1p-20f);
  for (int t = 0; t < m; t += 4) {
    const float4* __restrict__ p = XS + c + t;
    float4 a[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) a[u] = p[u];
    unsigned bits = 0;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const float rx = xi.x - a[u].x, ry = xi.y - a[u].y, rz = xi.z - a[u].z;
      const float d = __builtin_fmaf(rx, rx, __builtin_fmaf(ry, ry, rz * rz));
      bits |= (unsigned)((d < ct) & (t + u < m)) << u;
    }
    rel |= (unsigned long long)bits << t;  // t <= 60
  }
  return rel;
}

// visit_range over b..e-1 restricted to the in-range slots.
template <bool FAST, int G, class L, class F>
__device__ __forceinline__ void visit_range_c(int b, int e, float4 xi, float cut, const float4* __restrict__ XS,
                                              L&& ld, F&& f) {
  for (int c = b; c < e; c += 64) {
    unsigned long long rel = test_slots<FAST>(c, min(64, e - c), xi, cut, XS);
    for (;;) {
      int ks[G];
#pragma unroll
      for (int u = 0; u < G; ++u) {
        ks[u] = rel ? (int)__builtin_ctzll(rel) : -1;
        rel &= rel - 1;
      }
      if (ks[0] < 0) break;
      decltype(ld(0)) v[G];
#pragma unroll
      for (int u = 0; u < G; ++u) v[u] = ld(c + (ks[u] < 0 ? ks[0] : ks[u]));
#pragma unroll
      for (int u = 0; u < G; ++u) keep(v[u]);
#pragma unroll
      for (int u = 0; u < G; ++u)
        if (ks[u] >= 0) f(c + ks[u], v[u], 1);
    }
  }
}

// visit_all with the slot ranges restricted to their in-range candidates
// (r^2 < cut at xi): the same calls of f in the same order, minus calls every
// pass body returns from at once.  Classes holding runs keep the run walk,
// whose dependent run-start loads would not pay for a separate test phase.
template <bool FAST, int G, bool RUNS, int RG = 2, class L, class F>
__device__ __forceinline__ void visit_all_c(const Skin& K, bool rs, int l, int wx, int wxy,
                                            const int* __restrict__ cm, const int* __restrict__ cstart, int shift,
                                            float4 xi, float cut, const float4* __restrict__ XS, L&& ld, F&& f) {
#pragma unroll 1
  for (int r = 0; r < NPX_ROWS; ++r) {
    const int base = l + (r % 3 - 1) * wx + (r / 3 - 1) * wxy;
    const int c0 = cm[base - 1];
    const int c1 = cm[base];
    const int c2 = cm[base + 1];
    if (c1 == c0 + 1 && c2 == c0 + 2) {
      const int b = cstart[c0] - shift, e = cstart[c0 + 3] - shift;
      if (RUNS && rs && cls_runs(K, c0, 3))
        visit_runs<RG>(K, b, e, ld, f);
      else
        visit_range_c<FAST, G>(b, e, xi, cut, XS, ld, f);
    } else {
#pragma unroll 1
      for (int k = 0; k < 3; ++k) {
        const int c = (k == 0) ? c0 : (k == 1) ? c1 : c2;
        const int b = cstart[c] - shift, e = cstart[c + 1] - shift;
        if (RUNS && rs && cls_runs(K, c, 1))
          visit_runs<RG>(K, b, e, ld, f);
        else
          visit_range_c<FAST, G>(b, e, xi, cut, XS, ld, f);
      }
    }
  }
}

// Position of a candidate from what the pass's loader fetched.
template <class D>
__device__ __forceinline__ float4 cand_pos(const D& d) {
  if constexpr (std::is_same_v<D, float4>)
    return d;
  else
    return d.x;
}

// visit_all for a particle without a valid skin list, in two phases.  The
// test phase sweeps the 27 cells in the reference's order and appends every
// slot that passes the distance test (the superset test of test_slots; a run
// of k identical candidates appends its head k times) to the particle's
// in-range list; the visit phase then calls f over the list, G entries at a
// time, exactly as the reusing passes walk it.  f sees the same candidates in
// the same order as visit_all (k single calls where visit_all makes one call
// for a run of k, which adds the same floats: add_k is k additions in a row),
// minus calls it returns from at once.  The wave then evaluates its lanes'
// pair terms side by side in one loop, instead of one loop per row in which
// every lane waits for the lane with the most in-range candidates of that
// row.  The list is left for the reusing pass that follows.  Returns its
// length, or -1 when more than NPX_SKIN_CAP candidates pass: the test phase
// stops there and nothing is visited (the caller sweeps with visit_all_c).
#ifndef NPX_TEST_W
#define NPX_TEST_W 8  // candidates tested per round of the two-phase sweep (loads in flight per lane)
#endif
template <bool FAST, int G, bool RUNS, class L, class F>
__device__ __forceinline__ int visit_all_two(const Skin& K, bool rs, int q, float4 xi, float cut, int l, int wx,
                                             int wxy, const int* __restrict__ cm, const int* __restrict__ cstart,
                                             int shift, const float4* __restrict__ XS, L&& ld, F&& f) {
  const float ct = FAST ? cut : cut * (1.0f + 0x1p-20f);
  int e = 0;
  ListWriter out(K.ex, q);
  auto push = [&](int j) {
    out.push(j);
    ++e;
  };
  auto in = [&](const float4& a) {
    const float rx = xi.x - a.x, ry = xi.y - a.y, rz = xi.z - a.z;
    return __builtin_fmaf(rx, rx, __builtin_fmaf(ry, ry, rz * rz)) < ct;
  };
  // slots b .. en-1, NPX_TEST_W from one base address (up to NPX_TEST_W - 1
  // past en - 1: the x* buffers carry NPX_SLOT_PAD elements of padding), masked
  // (the tests of a round go to a bit mask without branches; only the
  // passing slots, about one in seven, take the append loop)
  auto range = [&](int b, int en) {
    for (int t = b; t < en && e <= NPX_SKIN_CAP; t += NPX_TEST_W) {
      const float4* __restrict__ p = XS + t;
      float4 a[NPX_TEST_W];
#pragma unroll
      for (int u = 0; u < NPX_TEST_W; ++u) a[u] = p[u];
      unsigned m = 0u;
#pragma unroll
      for (int u = 0; u < NPX_TEST_W; ++u) m |= (unsigned)in(a[u]) << u;
      if (en - t < NPX_TEST_W) m &= (1u << (en - t)) - 1u;
      while (m) {
        const int u = (int)__builtin_ctz(m);
        m &= m - 1u;
        push(t + u);
      }
    }
  };
  // runs of identical state in slots b .. en-1, two run heads per fetch
  auto runs = [&](int b, int en) {
    const int iend = K.pi[en - 1] + 1;
    for (int i = K.pi[b]; i < iend && e <= NPX_SKIN_CAP; i += 2) {
      const int j0 = K.ps[i];
      const int j1 = (i + 1 < iend) ? K.ps[i + 1] : en;
      const int j2 = (i + 2 < iend) ? K.ps[i + 2] : en;
      const float4 a0 = XS[j0], a1 = XS[(i + 1 < iend) ? j1 : j0];
      const bool i0 = in(a0), i1 = (i + 1 < iend) && in(a1);
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int j = h ? j1 : j0, k = (h ? j2 : j1) - j;
        if (h ? i1 : i0) {
          if (e + k <= NPX_SKIN_CAP) {
            for (int u = 0; u < k; ++u) push(j);
          } else {
            e += k;  // overflow: the list is dropped
            out.n += k;
          }
        }
      }
    }
  };
#pragma unroll 1
  for (int r = 0; r < NPX_ROWS && e <= NPX_SKIN_CAP; ++r) {
    const int base = l + (r % 3 - 1) * wx + (r / 3 - 1) * wxy;
    const int c0 = cm[base - 1];
    const int c1 = cm[base];
    const int c2 = cm[base + 1];
    if (c1 == c0 + 1 && c2 == c0 + 2) {
      const int b = cstart[c0] - shift, en = cstart[c0 + 3] - shift;
      if (b < en) {
        if (RUNS && rs && cls_runs(K, c0, 3))
          runs(b, en);
        else
          range(b, en);
      }
    } else {
#pragma unroll 1
      for (int k = 0; k < 3; ++k) {
        const int c = (k == 0) ? c0 : (k == 1) ? c1 : c2;
        const int b = cstart[c] - shift, en = cstart[c + 1] - shift;
        if (b < en) {
          if (RUNS && rs && cls_runs(K, c, 1))
            runs(b, en);
          else
            range(b, en);
        }
      }
    }
  }
  if (e > NPX_SKIN_CAP) return -1;
  out.flush();
  visit_list<G>(K.ex, q, e, ld, f);
  return e;
}

// acc += t, k times in a row (k equal terms of a run), bit for bit.
__device__ __forceinline__ void add_k(float& acc, float t, int k) { acc = (k == 1) ? acc + t : npx_rep_add(acc, t, k); }

// Testing pass (density, XSPH / curl): visits the in-range candidates of
// slot q at xi through the skin list, recording them for the pass that
// follows, or by the inline sweep, which records them when `record` (a
// reusing pass follows) and they fit (in-range count -1 otherwise).  C is the
// particle's candidate_disp.
template <bool FAST, int G, bool RUNS, class L, class F>
__device__ __forceinline__ void visit_test(const SceneC& S, const Skin& K, bool rs, int q, float C, float4 xi, float cut,
                                           const float4* __restrict__ XS, int l, int wxy,
                                           const int* __restrict__ cm, const int* __restrict__ cstart, int shift,
                                           bool record, int* __restrict__ st, L&& ld, F&& f) {
  const int n = (FAST || K.lstat) ? K.cnt[q] : -1;  // (no skin lists without the list policy)
  const bool held = K.cd[q] + C <= S.skin_allow;
  const int wv = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  if (!FAST && K.lstat && (wv & 7) == 0) {  // list statistics for the next step, sampled
    const unsigned long long tb = __ballot(1), hb = __ballot(held);
    if ((int)(threadIdx.x & 63) == (int)__builtin_ctzll(tb)) {
      int* c = st + 16 * ((wv >> 3) & (NPX_LST_COPIES - 1));
      atomicAdd(c, (int)__builtin_popcountll(tb));
      atomicAdd(c + 1, (int)__builtin_popcountll(hb));
    }
  }
  if (n >= 0 && held) {
    const unsigned long long rel = test_list<FAST>(K, q, n, xi, cut, XS);
    K.ecnt[q] = visit_rel<G>(K, q, rel, ld, f);
  } else {
    const int e = visit_all_two<FAST, G, RUNS>(K, rs, q, xi, cut, l, S.wx, wxy, cm, cstart, shift, XS, ld, f);
    if (e < 0) visit_all_c<FAST, G, RUNS>(K, rs, l, S.wx, wxy, cm, cstart, shift, xi, cut, XS, ld, f);
    K.ecnt[q] = record ? e : -1;
  }
}

// Reusing pass (delta-p, vorticity force): same positions as the testing
// pass before it, so its in-range list applies.
template <bool FAST, int G, bool RUNS, int RG = 2, int CG = G, class L, class F>
__device__ __forceinline__ void visit_reuse(const SceneC& S, const Skin& K, bool rs, int q, int l, int wxy,
                                            const int* __restrict__ cm, const int* __restrict__ cstart, int shift,
                                            float4 xi, float cut, const float4* __restrict__ XS, L&& ld, F&& f) {
  const int e = K.ecnt[q];
  if (e >= 0)
    visit_list<G>(K.ex, q, e, ld, f);
  else
    visit_all_c<FAST, CG, RUNS, RG>(K, rs, l, S.wx, wxy, cm, cstart, shift, xi, cut, XS, ld, f);
}

// ------------------------------------------------------------ pair terms
// The SPH terms of one candidate pair for each neighbour pass, shared by the
// per-lane kernels and the per-wave sweep of dense particles (below).  f(j,
// d, k, a) adds the pair's terms to the accumulators a[], k times over for a
// run of k identical candidates.  Every term is added with add_k, so with
// a[] at zero and k = 1 the call leaves the terms themselves in a[] (0 + t
// == t, and a +0 in place of a -0 term adds the same to a sum that is never
// -0).

struct XV {
  float4 x, v;
};
__device__ __forceinline__ void keep(const XV& d) {
  npx_keep(d.x);
  npx_keep(d.v);
}

struct XOW {
  float4 x;
  float vw, ow;
};
__device__ __forceinline__ void keep(const XOW& d) {
  npx_keep(make_float4(d.x.x, d.x.y, d.x.z, d.vw));
  __asm__ volatile("" ::"v"(d.ow));
}

// compute_density (pbf_solver.cpp:129-142) and compute_lambda (:144-179):
// a = {rho, sum |grad_j|^2, grad_i}.  k > 1: the k equal terms of a run of
// identical candidates (slot q can be one of them; its self term equals
// theirs: the density term, no gradient).
template <bool FAST>
struct DensityPair {
  float4 xi;
  int q;
  float h, h2, c6, cs, m, mr, sp_lo, sp_hi, cut;
  bool fd;
  __device__ __forceinline__ DensityPair(const SceneC& S, int q_, float4 xi_)
      : xi(xi_), q(q_), h(S.h), h2(S.h2), c6(S.c6), cs(S.cs), m(S.mass), mr(S.m_inv_rho0), sp_lo(S.sp_lo),
        sp_hi(S.sp_hi), cut(FAST ? S.h2 : S.cut2), fd(S.fdiv != 0) {}
  static constexpr int NT = 5;
  __device__ __forceinline__ void operator()(int j, const float4& xj, int k, float* a) const {
    const float rx = xi.x - xj.x, ry = xi.y - xj.y, rz = xi.z - xj.z;
    const float r_sq = dist2<FAST>(rx, ry, rz);
    if (r_sq >= cut) return;  // only reached on inline sweeps
    add_k(a[0], m * poly6(r_sq, h2, c6), k);
    if (j == q) return;
    bool in;
    const float co = FAST ? spiky_coeff_fast(r_sq, h, h2, cs, &in) : spiky_coeff_sq(r_sq, sp_lo, sp_hi, h, cs, fd, &in);
    if (!in) return;
    const float ax = mr * (co * rx), ay = mr * (co * ry), az = mr * (co * rz);
    add_k(a[1], dist2<FAST>(ax, ay, az), k);
    add_k(a[2], ax, k);
    add_k(a[3], ay, k);
    add_k(a[4], az, k);
  }
};

// compute_delta_position (pbf_solver.cpp:181-223) with s_corr: a = delta-p
// before the 1/rho0 factor.  FAST accumulates with FMA instead.
template <bool FAST>
struct DeltaPair {
  float4 xi;  // w: lambda_i
  int q;
  float h, h2, c6, cs, m, sp_lo, sp_hi, cut, wdq, negk;
  bool n4, scorr_on, fd;
  __device__ __forceinline__ DeltaPair(const SceneC& S, int q_, float4 xi_)
      : xi(xi_), q(q_), h(S.h), h2(S.h2), c6(S.c6), cs(S.cs), m(S.mass), sp_lo(S.sp_lo), sp_hi(S.sp_hi),
        cut(FAST ? S.h2 : S.cut2), wdq(S.wdq), negk(-S.corr_k), n4(S.corr_n == 4), scorr_on(S.wdq > 1e-10f),
        fd(S.fdiv != 0) {}
  static constexpr int NT = 3;
  __device__ __forceinline__ void operator()(int j, const float4& xj, int k, float* a) const {
    const float ls = xi.w + xj.w;
    if (j == q) return;
    const float rx = xi.x - xj.x, ry = xi.y - xj.y, rz = xi.z - xj.z;
    const float r_sq = dist2<FAST>(rx, ry, rz);
    if (r_sq >= cut) return;  // spiky_grad is zero: the term adds exactly 0
    bool in;
    const float co = FAST ? spiky_coeff_fast(r_sq, h, h2, cs, &in) : spiky_coeff_sq(r_sq, sp_lo, sp_hi, h, cs, fd, &in);
    if (!in) return;
    float sc = 0.0f;
    if (ls < -1e-10f && scorr_on) {
      float ratio = npx_div(poly6(r_sq, h2, c6), wdq, fd);
      ratio = (1.0f < ratio) ? 1.0f : ratio;  // std::min(ratio, 1.0f)
      float rp = ratio * ratio;
      if (n4) rp = rp * rp;
      sc = negk * rp;
    }
    const float coef = (ls + sc) * m;
    if (FAST) {
      const float cc = coef * co;
      a[0] = __builtin_fmaf(cc, rx, a[0]);
      a[1] = __builtin_fmaf(cc, ry, a[1]);
      a[2] = __builtin_fmaf(cc, rz, a[2]);
    } else {
      add_k(a[0], coef * (co * rx), k);
      add_k(a[1], coef * (co * ry), k);
      add_k(a[2], coef * (co * rz), k);
    }
  }
};

// apply_xsph_viscosity (pbf_solver.cpp:225-250) and the vorticity curl:
// a = {c_i, omega_i}.
struct OmegaPair {
  float4 xi, vi;
  int q;
  float h, h2, c6, cs, cut, sp_lo, sp_hi;
  bool xs, vort, fd;
  __device__ __forceinline__ OmegaPair(const SceneC& S, int q_, float4 xi_, float4 vi_)
      : xi(xi_), vi(vi_), q(q_), h(S.h), h2(S.h2), c6(S.c6), cs(S.cs), cut(S.cut2), sp_lo(S.sp_lo), sp_hi(S.sp_hi),
        xs(S.xsph_on != 0), vort(S.vort_on != 0), fd(S.fdiv != 0) {}
  static constexpr int NT = 6;
  __device__ __forceinline__ void operator()(int j, const XV& d, int k, float* a) const {
    if (j == q) return;
    const float rx = d.x.x - xi.x, ry = d.x.y - xi.y, rz = d.x.z - xi.z;
    const float r_sq = sqn3(rx, ry, rz);  // equals |x*_i - x*_j|^2 bit for bit
    if (r_sq >= cut) return;              // both terms are exact zeros beyond cut2
    const float4 vj = d.v;
    const float ax = vj.x - vi.x, ay = vj.y - vi.y, az = vj.z - vi.z;
    if (xs && r_sq < h2) {
      const float wv = poly6(r_sq, h2, c6);
      add_k(a[0], (vj.w * ax) * wv, k);
      add_k(a[1], (vj.w * ay) * wv, k);
      add_k(a[2], (vj.w * az) * wv, k);
    }
    if (vort) {
      bool in;
      const float co = spiky_coeff_sq(r_sq, sp_lo, sp_hi, h, cs, fd, &in);
      if (!in) return;
      const float gx = co * rx, gy = co * ry, gz = co * rz;
      add_k(a[3], vj.w * (ay * gz - az * gy), k);
      add_k(a[4], vj.w * (az * gx - ax * gz), k);
      add_k(a[5], vj.w * (ax * gy - ay * gx), k);
    }
  }
};

// Vorticity force gradient eta_i: a = eta_i.
struct VortPair {
  float4 xi;
  float oiw;  // |omega_i|
  int q;
  float h, cs, sp_lo, sp_hi;
  bool fd;
  __device__ __forceinline__ VortPair(const SceneC& S, int q_, float4 xi_, float oiw_)
      : xi(xi_), oiw(oiw_), q(q_), h(S.h), cs(S.cs), sp_lo(S.sp_lo), sp_hi(S.sp_hi), fd(S.fdiv != 0) {}
  static constexpr int NT = 3;
  __device__ __forceinline__ void operator()(int j, const XOW& d, int k, float* a) const {
    if (j == q) return;
    const float rx = xi.x - d.x.x, ry = xi.y - d.x.y, rz = xi.z - d.x.z;
    const float r_sq = sqn3(rx, ry, rz);
    bool in;
    const float co = spiky_coeff_sq(r_sq, sp_lo, sp_hi, h, cs, fd, &in);
    if (!in) return;
    const float sc = d.vw * (d.ow - oiw);
    add_k(a[0], sc * (co * rx), k);
    add_k(a[1], sc * (co * ry), k);
    add_k(a[2], sc * (co * rz), k);
  }
};

// ------------------------------------------------------------ per-particle ends

// lambda from the density sums (pbf_solver.cpp:144-179): 0 when C_i < 0,
// else -C_i / (sum + |grad_i|^2 + eps).
__device__ __forceinline__ void density_finish(const SceneC& S, int q, const float* a, float4* __restrict__ XS,
                                               float* __restrict__ RHO, float* __restrict__ LAM) {
  const float rho = a[0];
  RHO[q] = rho;
  const float Ci = rho * S.inv_rho0 - 1.0f;
  float lam = 0.0f;
  if (!(Ci < 0.0f)) {
    const float sgs = a[1] + sqn3(a[2], a[3], a[4]);
    lam = -Ci / (sgs + S.eps);
  }
  XS[q].w = lam;
  LAM[q] = lam;
}

// Velocity update with the h/dt speed clamp (pbf_solver.cpp:81-93); VT.w
// holds m / rho~ (rho~ = rho if > 1e-6 else rho0) for the XSPH sum.
__device__ __forceinline__ void velocity_update(const SceneC& S, int q, float px, float py, float pz,
                                                const float4* __restrict__ X, const float* __restrict__ RHO,
                                                float4* __restrict__ VT) {
  const float4 x = X[q];
  float vx = (px - x.x) * S.inv_dt, vy = (py - x.y) * S.inv_dt, vz = (pz - x.z) * S.inv_dt;
  const float sp = sqn3(vx, vy, vz);
  if (sp > S.max_speed_sq) {
    const float f = S.max_speed / sqrtf(sp);
    vx = vx * f;
    vy = vy * f;
    vz = vz * f;
  }
  const float r = RHO[q];
  const float rt = (r > 1e-6f) ? r : S.rho0;
  VT[q] = make_float4(vx, vy, vz, S.mass / rt);
}

// Jacobi apply (pbf_solver.cpp:71-77) into the other x* buffer, the velocity
// update on the last iteration, and the particle's summed displacement since
// the lists were built (returned; only the skin lists read it, so without
// them (track = false) it is not kept and 0 is returned).
__device__ __forceinline__ float delta_finish(const SceneC& S, const Skin& K, int q, float4 xi, const float* a,
                                              float4* __restrict__ XSn, int iter, bool last,
                                              const float4* __restrict__ X, const float* __restrict__ RHO,
                                              float4* __restrict__ VT, bool track = true) {
  const float px = xi.x + a[0] * S.inv_rho0, py = xi.y + a[1] * S.inv_rho0, pz = xi.z + a[2] * S.inv_rho0;
  XSn[q] = make_float4(px, py, pz, 0.0f);
  if (last) velocity_update(S, q, px, py, pz, X, RHO, VT);
  if (!track) return 0.0f;
  float moved = sqrtf(sqn3(px - xi.x, py - xi.y, pz - xi.z));
  moved = (iter == 0 ? 0.0f : K.cd[q]) + moved;
  K.cd[q] = moved;
  return moved;
}

// OM = (omega_i, |omega_i|); VX = v_i + xsph_c * c_i (v_i when XSPH is off).
__device__ __forceinline__ void omega_finish(const SceneC& S, int q, float4 vi, const float* a,
                                             float4* __restrict__ OM, float4* __restrict__ VX) {
  float vx = vi.x, vy = vi.y, vz = vi.z;
  if (S.vort_on) OM[q] = make_float4(a[3], a[4], a[5], sqrtf(sqn3(a[3], a[4], a[5])));
  if (S.xsph_on) {
    vx = vx + S.xsph_c * a[0];
    vy = vy + S.xsph_c * a[1];
    vz = vz + S.xsph_c * a[2];
  }
  VX[q] = make_float4(vx, vy, vz, 0.0f);
}

// Vorticity force from eta (when on), then the commit with the domain clamp
// (pbf_solver.cpp:102-126) into the new committed X / V.
__device__ __forceinline__ void commit_finish(const SceneC& S, int q, float4 xi, float4 vx4, float4 oi, const float* a,
                                              float4* __restrict__ X, float4* __restrict__ V) {
  float vx = vx4.x, vy = vx4.y, vz = vx4.z;
  if (S.vort_on) {
    const float el = sqrtf(sqn3(a[0], a[1], a[2]));
    if (el > 1e-9f) {
      const float inv = 1.0f / el;
      const float nx = a[0] * inv, ny = a[1] * inv, nz = a[2] * inv;
      const float fx = S.vort_eps * (ny * oi.z - nz * oi.y);
      const float fy = S.vort_eps * (nz * oi.x - nx * oi.z);
      const float fz = S.vort_eps * (nx * oi.y - ny * oi.x);
      vx = vx + S.dt * fx;
      vy = vy + S.dt * fy;
      vz = vz + S.dt * fz;
    }
  }
  float px = xi.x, py = xi.y, pz = xi.z;
  if (S.bounds) {
    if (px < S.clo0) { px = S.clo0; vx = 0.0f; } else if (px > S.chi0) { px = S.chi0; vx = 0.0f; }
    if (py < S.clo1) { py = S.clo1; vy = 0.0f; } else if (py > S.chi1) { py = S.chi1; vy = 0.0f; }
    if (pz < S.clo2) { pz = S.clo2; vz = 0.0f; } else if (pz > S.chi2) { pz = S.chi2; vz = 0.0f; }
  }
  X[q] = make_float4(px, py, pz, 0.0f);
  V[q] = make_float4(vx, vy, vz, 0.0f);
}

// ------------------------------------------------------------ dense particles
// A particle whose skin list overflows (more than NPX_SKIN_CAP candidates
// within the skin radius: a dense sheet or the surroundings of a stack) in
// crowded cells sweeps thousands of candidates in every pass.  A lane per particle would
// leave those sweeps to the few waves that hold such particles, each walking
// its candidates one dependent fetch after another.  Instead the build pass
// marks them (heads only, exact FP32 path only) and groups them: the marked
// particles of one 64-slot window that share a cell, and so share the
// candidate sequence.  It lists one leader per batch of NPX_DENSE_BATCH
// members of a group.  The per-lane kernels skip them, and one wave per
// batch sweeps the candidates 64 at a time for all its members at once:
// lane t fetches candidate t of the chunk
// and evaluates its pair terms against each member (a[] from zero, k = 1);
// the in-range terms go to LDS compacted in candidate order, one row per
// (member, accumulator); and one lane per row adds its row in order, k
// times for a run of k, as the per-lane sweep does.  Sums, order and results
// are those of the per-lane sweep bit for bit.

// Particles whose skin list overflows and whose 27 cells hold more than
// this many candidates go on the dense list; sparser overflows keep the
// per-lane inline sweep.
#define NPX_DENSE_MIN 256

// Skin-list count of a listed dense particle (other overflows count -1).
#define NPX_CNT_DENSE (-2)

// Members of a dense group swept together by one wave.
#ifndef NPX_DENSE_BATCH
#define NPX_DENSE_BATCH 2
#endif
// Waves of a neighbour pass that sweep the dense list beside its per-lane
// blocks (the build pass's separate dense launch uses up to 8192).
#ifndef NPX_DENSE_WAVES
#define NPX_DENSE_WAVES 32768
#endif
// LDS row stride of a dense sweep: rows start 4 banks apart.
#define NPX_DENSE_RS 68
static_assert(NPX_DENSE_RS % 4 == 0 && NPX_DENSE_RS >= 64 + 4, "16-byte rows with a four-entry tail");

// LDS floats per wave of a dense sweep with NT accumulators: the term rows,
// their run counts and the per-member term counts.
template <int NT>
__host__ __device__ constexpr int npx_dense_lds_floats() {
  return (NPX_DENSE_BATCH * NT * NPX_DENSE_RS + NPX_DENSE_BATCH * NPX_DENSE_RS + NPX_DENSE_BATCH + 3) & ~3;
}

// Orders one wave's LDS writes before its following reads (the dense
// sweeps' waves work independently, so a block barrier does not fit).
__device__ __forceinline__ void npx_wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Whether slot q (a head of a scene with solver iterations, after this
// step's build pass) is on the dense list.
__device__ __forceinline__ bool is_dense(const SceneC& S, const Skin& K, int q) {
  return K.dense && S.iters > 0 && K.cnt[q] == NPX_CNT_DENSE;
}

// Candidates in the 27 cells of window cell l (duplicates counted).
__device__ __forceinline__ int candidate_total(int l, int wx, int wxy, const int* __restrict__ cm,
                                               const int* __restrict__ cstart) {
  int total = 0;
#pragma unroll 1
  for (int r = 0; r < NPX_ROWS; ++r) {
    const int base = l + (r % 3 - 1) * wx + (r / 3 - 1) * wxy;
    const int c0 = cm[base - 1], c1 = cm[base], c2 = cm[base + 1];
    if (c1 == c0 + 1 && c2 == c0 + 2)
      total += cstart[c0 + 3] - cstart[c0];
    else
      total += (cstart[c0 + 1] - cstart[c0]) + (cstart[c1 + 1] - cstart[c1]) + (cstart[c2 + 1] - cstart[c2]);
  }
  return total;
}

// Lists q if it leads a batch: the calling lanes of the wave with window
// cell l form a group, and every NPX_DENSE_BATCH-th of them in lane order
// (from the lowest) leads the batch of itself and the next ones (one atomic
// per wave).
__device__ __forceinline__ void dense_push(const Skin& K, int q, int l) {
  const int lane = (int)(threadIdx.x & 63);
  unsigned long long todo = __ballot(1);
  bool lead = false;
  while (todo) {
    const int first = (int)__builtin_ctzll(todo);
    const int lc = __shfl(l, first, 64);
    const unsigned long long same = __ballot(l == lc) & todo;
    if (l == lc) lead = ((int)__builtin_popcountll(same & ((1ull << lane) - 1ull)) % NPX_DENSE_BATCH) == 0;
    todo &= ~same;
  }
  const unsigned long long m = __ballot(lead);
  const int leader = (int)__builtin_ctzll(m);
  const int rank = (int)__builtin_popcountll(m & ((1ull << lane) - 1ull));
  int base = 0;
  if (lane == leader) base = atomicAdd(K.dn, (int)__builtin_popcountll(m));
  base = __shfl(base, leader, 64);
  if (lead) K.dl[base + rank] = q;
}

// The dense batch of leader q0 (window cell l): the marked heads of q0's
// 64-slot window with cell l, from q0 on, at most NPX_DENSE_BATCH of them.
// Their candidates are swept in the reference's order, as visit_all walks
// them, by the whole 64-lane block; make(q) gives member q's pair functor
// (its position in .xi), and fin(q, a) runs on member q's own lane with its
// NT sums.  Each 64-candidate chunk writes, per member, the terms of its
// in-range candidates compacted in candidate order into one padded LDS row
// per accumulator; lane p * NT + u then adds row (p, u) in order.  In a
// range without runs it reads four entries per 16-byte LDS read: each row's
// tail after its last term holds four -0.0f, and x + -0.0f is x for every
// float x.
// sel >= 0: only member sel of the batch (another wave takes the others).
template <int NT, class L, class MK, class FIN>
__device__ __forceinline__ void dense_group(const SceneC& S, const Skin& K, int q0, bool rs, int l, int wxy,
                                            const int* __restrict__ cm, const int* __restrict__ cstart, int shift,
                                            const int* __restrict__ cell, float cut, float* __restrict__ lds, L&& ld,
                                            MK&& make, FIN&& fin, int sel = -1) {
  constexpr int GB = NPX_DENSE_BATCH;
  constexpr int RS = NPX_DENSE_RS;
  static_assert(GB * NT <= 64, "one summing lane per row");
  float* __restrict__ tl = lds;                                  // [GB * NT][RS]
  int* __restrict__ kl = reinterpret_cast<int*>(lds + GB * NT * RS);  // [GB][RS]
  int* __restrict__ cl = kl + GB * RS;                           // [GB]
  const int t = (int)(threadIdx.x & 63);
  const int wb = q0 & ~63;
  const int qt = wb + t;
  const bool member = qt < S.seg_off + S.n && (!rs || K.rstart[qt]) && K.cnt[qt] == NPX_CNT_DENSE && cell[qt] == l;
  unsigned long long left = __ballot(member) & ~((1ull << (q0 & 63)) - 1ull);
  const int ps = t / NT, us = t - ps * NT;  // the row this lane adds
  {
    int qs[GB];
    int nb = 0;
    unsigned long long bm = 0ull;
#pragma unroll
    for (int p = 0; p < GB; ++p) {
      qs[p] = wb;
      if (left) {
        const int b = (int)__builtin_ctzll(left);
        qs[p] = wb + b;
        if (sel < 0 || sel == p) bm |= 1ull << b;
        left &= left - 1ull;
        ++nb;
      }
    }
    if (sel >= 0) {  // uniform
      if (sel >= nb) return;
#pragma unroll
      for (int p = 1; p < GB; ++p) qs[0] = (p == sel) ? qs[p] : qs[0];
      nb = 1;
    }
    float acc = 0.0f;
    using D = decltype(ld(0));
    // lane t's candidate of a chunk: m <= 64 slots from c, or m runs from
    // piece c (runs: iend is the range's end piece and e its end slot; k the
    // run's length)
    auto cand = [&](int c, int m, bool runs, int iend, int e, int& k) -> int {
      const int i = c + (t < m ? t : 0);
      k = 1;
      if (!runs) return i;
      const int j = K.ps[i];
      k = ((i + 1 < iend) ? K.ps[i + 1] : e) - j;
      return j;
    };
    // one chunk of m candidates, lane t's candidate j (run length k, data d)
    auto chunk = [&](int m, int j, int k, const D& d, bool runs) {
      const bool valid = t < m;
      const float4 xj = cand_pos(d);
#pragma unroll
      for (int p = 0; p < GB; ++p) {
        if (p < nb) {
          const auto P = make(qs[p]);
          const bool in = valid && dist2<false>(P.xi.x - xj.x, P.xi.y - xj.y, P.xi.z - xj.z) < cut;
          float a[NT];
#pragma unroll
          for (int u = 0; u < NT; ++u) a[u] = 0.0f;
          if (in) P(j, d, 1, a);
          const unsigned long long mask = __ballot(in);
          const int cnt = (int)__builtin_popcountll(mask);
          if (in) {
            const int pos = (int)__builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32),
                                                           __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
#pragma unroll
            for (int u = 0; u < NT; ++u) tl[(p * NT + u) * RS + pos] = a[u];
            kl[p * RS + pos] = k;
          }
          if (t >= cnt && t < cnt + 4) {  // the rows' tails (RS >= 68 entries)
#pragma unroll
            for (int u = 0; u < NT; ++u) tl[(p * NT + u) * RS + t] = -0.0f;
          }
          if (t == 0) cl[p] = cnt;
        }
      }
      npx_wave_lds_sync();
      if (ps < nb) {
        const float* row = tl + (ps * NT + us) * RS;
        const int* kr = kl + ps * RS;
        const int n = cl[ps];
        if (runs) {
          for (int i = 0; i < n; ++i) add_k(acc, row[i], kr[i]);
        } else {  // no runs in the range: every count is 1
          for (int i = 0; i < n; i += 4) {
            const float4 v = *reinterpret_cast<const float4*>(row + i);
            acc = acc + v.x;
            acc = acc + v.y;
            acc = acc + v.z;
            acc = acc + v.w;
          }
        }
      }
      npx_wave_lds_sync();
    };
    // a range's chunks in order
    auto range = [&](int b, int e, bool runs) {
      if (b >= e) return;
      const int c0 = runs ? K.pi[b] : b;
      const int cend = runs ? K.pi[e - 1] + 1 : e;
      for (int c = c0; c < cend; c += 64) {
        int k;
        const int j = cand(c, min(64, cend - c), runs, cend, e, k);
        chunk(min(64, cend - c), j, k, ld(j), runs);
      }
    };
    for (int r = 0; r < NPX_ROWS; ++r) {
      const int base = l + (r % 3 - 1) * S.wx + (r / 3 - 1) * wxy;
      const int c0 = cm[base - 1];
      const int c1 = cm[base];
      const int c2 = cm[base + 1];
      if (c1 == c0 + 1 && c2 == c0 + 2) {
        range(cstart[c0] - shift, cstart[c0 + 3] - shift, rs && cls_runs(K, c0, 3));
      } else {
        for (int u = 0; u < 3; ++u) {
          const int c = (u == 0) ? c0 : (u == 1) ? c1 : c2;
          range(cstart[c] - shift, cstart[c + 1] - shift, rs && cls_runs(K, c, 1));
        }
      }
    }
    // member slot qt has rank r in the batch: its sums are in lanes r*NT..
    const bool mine = (bm >> t) & 1ull;
    // (the count of bm's bits below lane t, without a 64-bit lane mask that
    // would stay live across the sweep)
    const int r = (int)__builtin_amdgcn_mbcnt_hi((unsigned)(bm >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)bm, 0u));
    const int src = mine ? r * NT : 0;
    float a[NT];
#pragma unroll
    for (int u = 0; u < NT; ++u) a[u] = __shfl(acc, src + u, 64);
    if (mine) fin(qt, a);
  }
}

// Scene data of dense group leader q (uniform over the block).
#define NPX_DENSE_PROLOGUE()                                 \
  const int q = K.dl[w];                                     \
  const int s = blk_scene[q / NPX_BLOCK];                    \
  const SceneC& S = scenes[s];                               \
  const int shift = cstart[S.cls_base] - S.seg_off;          \
  const int wxy = S.wx * S.wy;                               \
  const int* __restrict__ cm = clsmap + S.cls_base;          \
  const int l = cell[q];                                     \
  const bool rs = K.runs_s[s] != 0;

// The dense list's share of the neighbour passes (defined with each pass).
// The passes after the build run it in their first bp.nd launch blocks,
// beside the per-lane blocks, one wave per batch.
__device__ __forceinline__ void dense_density(int, int, const SceneC* __restrict__, const int* __restrict__,
                                              const int* __restrict__, const int* __restrict__,
                                              const int* __restrict__, float4* __restrict__, float* __restrict__,
                                              float* __restrict__, int, const Skin&, bool split = false);
__device__ __forceinline__ void dense_delta(int, int, const SceneC* __restrict__, const int* __restrict__,
                                            const int* __restrict__, const int* __restrict__, const int* __restrict__,
                                            const float4* __restrict__, float4* __restrict__, int,
                                            const float4* __restrict__, const float* __restrict__,
                                            float4* __restrict__, const Skin&);
__device__ __forceinline__ void dense_omega(int, int, const SceneC* __restrict__, const int* __restrict__,
                                            const int* __restrict__, const int* __restrict__, const int* __restrict__,
                                            const float4* __restrict__, const float4* __restrict__,
                                            float4* __restrict__, float4* __restrict__, const Skin&);
__device__ __forceinline__ void dense_commit(int, int, const SceneC* __restrict__, const int* __restrict__,
                                             const int* __restrict__, const int* __restrict__, const int* __restrict__,
                                             const float4* __restrict__, const float4* __restrict__,
                                             const float4* __restrict__, const float4* __restrict__,
                                             float4* __restrict__, float4* __restrict__, const Skin&);

// Dense-list batch of wave (threadIdx.x / 64) of dense launch block blockIdx.x,
// and the stride over the batches.
#define NPX_DENSE_WAVE() ((int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)))
#define NPX_DENSE_STRIDE(bp) ((bp).nd * (int)(blockDim.x >> 6))

// compute_density (pbf_solver.cpp:129-142) fused with compute_lambda
// (:144-179): rho_i sums m W over all in-range candidates (self and
// duplicates included); for j != i, sum |grad_j|^2 and grad_i; lambda is 0
// when C_i < 0, else -C_i / (sum + |grad_i|^2 + eps).  BUILD (the first
// iteration) lists the crowded particles for the dense kernels and, under
// list policy 1, builds the step's skin lists; later iterations test the
// lists where they exist.  The build pass (BUILD) is then two launches: with
// LISTS it serves the scenes whose flag says this step builds skin lists,
// without LISTS the others (every block belongs to one scene and leaves at
// once when its scene is the other launch's); under policy 0 only the
// list-free launch is made.  The list-free variant carries no list-building
// code and runs at the testing passes' occupancy.
template <bool FAST, bool BUILD, bool LISTS = true, bool PERSIST = false>
__global__ void __launch_bounds__(NPX_MAX_BLOCK) __attribute__((amdgpu_waves_per_eu((BUILD && LISTS) ? 4 : 6)))
k_density_lambda(const SceneC* __restrict__ scenes, const int* __restrict__ blk_scene, const int* __restrict__ clsmap,
                 const int* __restrict__ cstart, const int* __restrict__ cell, float4* __restrict__ XS,
                 float* __restrict__ RHO, float* __restrict__ LAM, int iter, Skin K, BlockPlan bp) {
  if (!FAST && !BUILD && (int)blockIdx.x < bp.nd) {
    dense_density(NPX_DENSE_WAVE(), NPX_DENSE_STRIDE(bp), scenes, blk_scene, clsmap, cstart, cell, XS, RHO, LAM, iter,
                  K);
    return;
  }
  npx_for_blocks<PERSIST>(bp, [&](const int lb) {
    if (BUILD && !FAST) {
      // the scene's flag picks the build launch (FAST always builds lists); tested
      // from the block's first slot (one scene per block) before the prologue's loads
      const int s0 = blk_scene[lb * (int)blockDim.x / NPX_BLOCK];
      if (s0 < 0 || (K.lst[NPX_LST_STRIDE * s0 + NPX_LST_BUILD] != 0) != LISTS) return;
    }
    NPX_NBR_PROLOGUE(lb);
    const bool rs = K.runs_s[s] != 0;
    if (iter >= S.iters || !active || (rs && !K.rstart[q])) return;  // run members take their head's values
    if (!BUILD && !FAST && is_dense(S, K, q)) return;                  // the dense kernel's particle
    if (BUILD && !FAST && K.dense && candidate_total(l, S.wx, wxy, cm, cstart) > NPX_DENSE_MIN) {
      K.cnt[q] = NPX_CNT_DENSE;  // crowded: the per-wave dense sweep takes it, no list needed
      K.ecnt[q] = -1;
      dense_push(K, q, l);
      return;
    }
    const float4 xi = XS[q];
    const DensityPair<FAST> P(S, q, xi);
    const float cut = P.cut;
    float a[5] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    auto ld = [&](int j) { return XS[j]; };
    auto body = [&](int j, float4 xj, int k) { P(j, xj, k, a); };
    if (BUILD) {
      int n = -1;
      unsigned long long rel = 0;
      if (LISTS) {  // this step builds skin lists
        Rows w;
        if (sweep_rows<FAST>(w, K, rs, xi, cut, S.skin_cut2, XS, l, S.wx, wxy, cm, cstart, shift))
          n = store_skin(w, K, q, &rel);
        else
          n = sweep_direct<FAST, !FAST>(K, rs, q, xi, cut, S.skin_cut2, XS, l, S.wx, wxy, cm, cstart, shift, &rel);
        if (n > NPX_SKIN_CAP) n = -1;
      }
      K.cnt[q] = n;
      if (n >= 0) {
        K.ecnt[q] = visit_rel<4>(K, q, rel, ld, body);
      } else {
        const int e = visit_all_two<FAST, 4, !FAST>(K, rs, q, xi, cut, l, S.wx, wxy, cm, cstart, shift, XS, ld, body);
        if (e < 0) visit_all_c<FAST, 4, !FAST>(K, rs, l, S.wx, wxy, cm, cstart, shift, xi, cut, XS, ld, body);
        K.ecnt[q] = e;
      }
    } else {
      // the list bound matters only where lists exist or their statistics
      // are sampled (list policy 1, or the FAST path, which always builds them)
      const float C = (FAST || K.lstat) ? candidate_disp(K, l, S.wx, wxy, cm) : 0.0f;
      visit_test<FAST, 4, !FAST>(S, K, rs, q, C, xi, cut, XS, l, wxy, cm, cstart, shift, true, K.lst + NPX_LST_STRIDE * s, ld, body);
    }
    // the slot's addresses are formed again here rather than kept live across
    // the sweep (an empty asm hides q's value from the compiler)
    int qf = q;
    __asm__ volatile("" : "+v"(qf));
    density_finish(S, qf, a, XS, RHO, LAM);
  });
}

// k_density_lambda for the dense list, batches w0, w0 + ws, ... (iterations
// of the exact path).  split: w counts (batch, member) pairs, one member of
// a batch per wave.
__device__ __forceinline__ void dense_density(int w0, int ws, const SceneC* __restrict__ scenes,
                                              const int* __restrict__ blk_scene, const int* __restrict__ clsmap,
                                              const int* __restrict__ cstart, const int* __restrict__ cell,
                                              float4* __restrict__ XS, float* __restrict__ RHO,
                                              float* __restrict__ LAM, int iter, const Skin& K, bool split) {
  extern __shared__ float4 npx_dyn4[];  // 16-byte aligned: the dense rows are read as float4
  float* npx_dyn = reinterpret_cast<float*>(npx_dyn4);
  const int sh = split ? (NPX_DENSE_BATCH == 2 ? 1 : 0) : 0;  // (batch, member) pairs when splitting
  for (int wm = w0; wm < (*K.dn << sh); wm += ws) {
    const int w = wm >> sh;
    NPX_DENSE_PROLOGUE();
    if (iter >= S.iters) continue;
    dense_group<5>(S, K, q, rs, l, wxy, cm, cstart, shift, cell, S.cut2, npx_dyn + (threadIdx.x >> 6) * npx_dense_lds_floats<5>(), [&](int j) { return XS[j]; },
                   [&](int qm) { return DensityPair<false>(S, qm, XS[qm]); },
                   [&](int qm, const float* a) { density_finish(S, qm, a, XS, RHO, LAM); }, sh ? (wm & 1) : -1);
  }
}

// The build pass's dense list (the first iteration), one wave per block and
// one member of a batch per wave: the launch runs alone and its end is set
// by its longest sweeps, which one member per wave shortens.
__global__ void __launch_bounds__(64)
k_density_dense(const SceneC* __restrict__ scenes, const int* __restrict__ blk_scene, const int* __restrict__ clsmap,
                const int* __restrict__ cstart, const int* __restrict__ cell, float4* __restrict__ XS,
                float* __restrict__ RHO, float* __restrict__ LAM, int iter, Skin K) {
  dense_density((int)blockIdx.x, (int)gridDim.x, scenes, blk_scene, clsmap, cstart, cell, XS, RHO, LAM, iter, K,
                true);
}

// compute_delta_position (pbf_solver.cpp:181-223) with s_corr, fused with the
// Jacobi apply (:71-77) into the other x* buffer.  On the last iteration it
// also performs the velocity update with the h/dt speed clamp (:81-93).
// Adds the particle's |delta x*| to its summed displacement and keeps the
// largest sum of each class for the skin-list test.
template <bool FAST, bool PERSIST = false>
__global__ void __launch_bounds__(NPX_MAX_BLOCK) __attribute__((amdgpu_waves_per_eu(7)))
k_delta(const SceneC* __restrict__ scenes, const int* __restrict__ blk_scene, const int* __restrict__ clsmap,
        const int* __restrict__ cstart, const int* __restrict__ cell, const float4* __restrict__ XS,
        float4* __restrict__ XSn, int iter, const float4* __restrict__ X, const float* __restrict__ RHO,
        float4* __restrict__ VT, Skin K, BlockPlan bp) {
  if (!FAST && (int)blockIdx.x < bp.nd) {
    dense_delta(NPX_DENSE_WAVE(), NPX_DENSE_STRIDE(bp), scenes, blk_scene, clsmap, cstart, cell, XS, XSn, iter, X, RHO,
                VT, K);
    return;
  }
  npx_for_blocks<PERSIST>(bp, [&](const int lb) {
    NPX_NBR_PROLOGUE(lb);
    const bool last = (iter == max(S.iters, 1) - 1);
    if (iter >= S.iters) {
      // scenes with fewer iterations than the batch maximum pass x* through so
      // every scene ends in the same ping-pong buffer
      if (!active) return;
      const float4 xi = XS[q];
      XSn[q] = xi;
      if (last) velocity_update(S, q, xi.x, xi.y, xi.z, X, RHO, VT);
      return;
    }
    float moved = 0.0f;
    const bool rs = K.runs_s[s] != 0;
    // run members take their head's values; dense-list particles are the dense kernel's
    if (active && !(rs && !K.rstart[q]) && !(!FAST && is_dense(S, K, q))) {
      const float4 xi = XS[q];
      const DeltaPair<FAST> P(S, q, xi);
      float a[3] = {0.0f, 0.0f, 0.0f};
      auto ld = [&](int j) { return XS[j]; };
      auto body = [&](int j, float4 xj, int k) { P(j, xj, k, a); };  // k > 1 only without FAST
      // one run / two slots at a time on the inline sweep: within 72 VGPRs
      visit_reuse<FAST, 4, !FAST, 1, 2>(S, K, rs, q, l, wxy, cm, cstart, shift, xi, P.cut, XS, ld, body);
      moved = delta_finish(S, K, q, xi, a, XSn, iter, last, X, RHO, VT, FAST || K.lstat);
    }
    if (!FAST && !K.lstat) return;  // displacements feed only the skin lists' bound
    // Per-class maxima: slots are sorted by class, so the lanes of a class form
    // one run in the wave; a segmented max leaves the run's maximum in its first
    // lane, which alone does the atomic.  (The own class of a particle is the
    // class of its cell, cm[l].)
    const int cls = active ? cm[l] : -1 - (int)(threadIdx.x & 63);
    const int lane = (int)(threadIdx.x & 63);
  #pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const float mo = __shfl_down(moved, o, 64);
      const int co = __shfl_down(cls, o, 64);
      if (lane + o < 64 && co == cls) moved = fmaxf(moved, mo);
    }
    const int prev = __shfl_up(cls, 1, 64);
    if (active && (lane == 0 || prev != cls) && moved > 0.0f) atomicMax(K.cmax + cls, __float_as_uint(moved));
  });
}

// k_delta for the dense list, batches w0, w0 + ws, ...
__device__ __forceinline__ void dense_delta(int w0, int ws, const SceneC* __restrict__ scenes,
                                            const int* __restrict__ blk_scene, const int* __restrict__ clsmap,
                                            const int* __restrict__ cstart, const int* __restrict__ cell,
                                            const float4* __restrict__ XS, float4* __restrict__ XSn, int iter,
                                            const float4* __restrict__ X, const float* __restrict__ RHO,
                                            float4* __restrict__ VT, const Skin& K) {
  extern __shared__ float4 npx_dyn4[];  // 16-byte aligned: the dense rows are read as float4
  float* npx_dyn = reinterpret_cast<float*>(npx_dyn4);
  for (int w = w0; w < *K.dn; w += ws) {
    NPX_DENSE_PROLOGUE();
    if (iter >= S.iters) continue;  // the pass-through is the per-lane kernel's
    const bool last = (iter == max(S.iters, 1) - 1);
    dense_group<3>(S, K, q, rs, l, wxy, cm, cstart, shift, cell, S.cut2, npx_dyn + (threadIdx.x >> 6) * npx_dense_lds_floats<3>(), [&](int j) { return XS[j]; },
                   [&](int qm) { return DeltaPair<false>(S, qm, XS[qm]); },
                   [&](int qm, const float* a) {
                     const float moved = delta_finish(S, K, qm, XS[qm], a, XSn, iter, last, X, RHO, VT, K.lstat != 0);
                     if (moved > 0.0f) atomicMax(K.cmax + cm[l], __float_as_uint(moved));
                   });
  }
}

// apply_xsph_viscosity (pbf_solver.cpp:225-250) fused with the vorticity
// curl (this project's definition; the reference declares vorticity_epsilon
// but never uses it): one testing pass at the final x* gives
//   c_i     = sum_{j!=i, r^2<h^2} (m/rho~_j) (v_j - v_i) W(x*_i - x*_j)
//   omega_i = sum_{j!=i} (m/rho~_j) (v_j - v_i) x gradW(x*_j - x*_i)
// VX = v_i + xsph_c * c_i (v_i when XSPH is off); OM = (omega_i, |omega_i|).
template <bool PERSIST = false>
__global__ void __launch_bounds__(NPX_MAX_BLOCK)
k_omega_xsph(const SceneC* __restrict__ scenes, const int* __restrict__ blk_scene, const int* __restrict__ clsmap,
             const int* __restrict__ cstart, const int* __restrict__ cell, const float4* __restrict__ XS,
             const float4* __restrict__ VT, float4* __restrict__ OM, float4* __restrict__ VX, Skin K, BlockPlan bp) {
  if ((int)blockIdx.x < bp.nd) {
    dense_omega(NPX_DENSE_WAVE(), NPX_DENSE_STRIDE(bp), scenes, blk_scene, clsmap, cstart, cell, XS, VT, OM, VX, K);
    return;
  }
  npx_for_blocks<PERSIST>(bp, [&](const int lb) {
    NPX_NBR_PROLOGUE(lb);
    const bool rs = K.runs_s[s] != 0;
    if (!active || (rs && !K.rstart[q])) return;  // run members take their head's values
    const float4 vi = VT[q];
    const bool xs = S.xsph_on != 0, vort = S.vort_on != 0;
    float a[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    if (xs || vort) {
      if (is_dense(S, K, q)) return;  // the dense kernel's particle
      const float4 xi = XS[q];
      const OmegaPair P(S, q, xi, vi);
      auto ld = [&](int j) { return XV{XS[j], VT[j]}; };
      auto body = [&](int j, const XV& d, int k) { P(j, d, k, a); };
      if (S.iters > 0) {  // the skin lists exist once a density pass has run
        const float C = K.lstat ? candidate_disp(K, l, S.wx, wxy, cm) : 0.0f;
        // the vorticity force pass reuses the in-range list
        visit_test<false, 4, true>(S, K, rs, q, C, xi, P.cut, XS, l, wxy, cm, cstart, shift, vort, K.lst + NPX_LST_STRIDE * s, ld,
                                   body);
      } else {
        K.ecnt[q] = -1;
        visit_all<2, false>(K, false, l, S.wx, wxy, cm, cstart, shift, ld, body);
      }
    }
    omega_finish(S, q, vi, a, OM, VX);
  });
}

// k_omega_xsph for the dense list, batches w0, w0 + ws, ...
__device__ __forceinline__ void dense_omega(int w0, int ws, const SceneC* __restrict__ scenes,
                                            const int* __restrict__ blk_scene, const int* __restrict__ clsmap,
                                            const int* __restrict__ cstart, const int* __restrict__ cell,
                                            const float4* __restrict__ XS, const float4* __restrict__ VT,
                                            float4* __restrict__ OM, float4* __restrict__ VX, const Skin& K) {
  extern __shared__ float4 npx_dyn4[];  // 16-byte aligned: the dense rows are read as float4
  float* npx_dyn = reinterpret_cast<float*>(npx_dyn4);
  for (int w = w0; w < *K.dn; w += ws) {
    NPX_DENSE_PROLOGUE();
    if (!S.xsph_on && !S.vort_on) continue;  // no neighbour sum: the per-lane kernel's
    dense_group<6>(S, K, q, rs, l, wxy, cm, cstart, shift, cell, S.cut2, npx_dyn + (threadIdx.x >> 6) * npx_dense_lds_floats<6>(),
                   [&](int j) { return XV{XS[j], VT[j]}; },
                   [&](int qm) { return OmegaPair(S, qm, XS[qm], VT[qm]); },
                   [&](int qm, const float* a) { omega_finish(S, qm, VT[qm], a, OM, VX); });
  }
}

// Vorticity force (N = eta / |eta|, eta_i = sum_{j!=i} (m/rho~_j)
// (|omega_j| - |omega_i|) gradW(x*_i - x*_j); v += dt * eps (N x omega_i))
// after XSPH, then the commit with the domain clamp (pbf_solver.cpp:102-126)
// into the new committed X / V.
template <bool PERSIST = false>
__global__ void __launch_bounds__(NPX_MAX_BLOCK)
k_vort_commit(const SceneC* __restrict__ scenes, const int* __restrict__ blk_scene, const int* __restrict__ clsmap,
              const int* __restrict__ cstart, const int* __restrict__ cell, const float4* __restrict__ XS,
              const float4* __restrict__ VT, const float4* __restrict__ OM, const float4* __restrict__ VX,
              float4* __restrict__ X, float4* __restrict__ V, Skin K, BlockPlan bp) {
  if ((int)blockIdx.x < bp.nd) {
    dense_commit(NPX_DENSE_WAVE(), NPX_DENSE_STRIDE(bp), scenes, blk_scene, clsmap, cstart, cell, XS, VT, OM, VX, X, V,
                 K);
    return;
  }
  npx_for_blocks<PERSIST>(bp, [&](const int lb) {
    NPX_NBR_PROLOGUE(lb);
    const bool rs = K.runs_s[s] != 0;
    if (!active || (rs && !K.rstart[q])) return;  // run members take their head's values
    const float4 xi = XS[q];
    float a[3] = {0.0f, 0.0f, 0.0f};
    float4 oi = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (S.vort_on) {
      if (is_dense(S, K, q)) return;  // the dense kernel's particle
      oi = OM[q];
      const VortPair P(S, q, xi, oi.w);
      auto ld = [&](int j) { return XOW{XS[j], VT[j].w, OM[j].w}; };
      auto body = [&](int j, const XOW& d, int k) { P(j, d, k, a); };
      visit_reuse<false, 4, true>(S, K, rs, q, l, wxy, cm, cstart, shift, xi, S.cut2, XS, ld, body);  // list of the XSPH / curl pass
    }
    commit_finish(S, q, xi, VX[q], oi, a, X, V);
  });
}

// k_vort_commit for the dense list (scenes with vorticity), batches w0,
// w0 + ws, ...
__device__ __forceinline__ void dense_commit(int w0, int ws, const SceneC* __restrict__ scenes,
                                             const int* __restrict__ blk_scene, const int* __restrict__ clsmap,
                                             const int* __restrict__ cstart, const int* __restrict__ cell,
                                             const float4* __restrict__ XS, const float4* __restrict__ VT,
                                             const float4* __restrict__ OM, const float4* __restrict__ VX,
                                             float4* __restrict__ X, float4* __restrict__ V, const Skin& K) {
  extern __shared__ float4 npx_dyn4[];  // 16-byte aligned: the dense rows are read as float4
  float* npx_dyn = reinterpret_cast<float*>(npx_dyn4);
  for (int w = w0; w < *K.dn; w += ws) {
    NPX_DENSE_PROLOGUE();
    if (!S.vort_on) continue;  // commit only: the per-lane kernel's
    dense_group<3>(S, K, q, rs, l, wxy, cm, cstart, shift, cell, S.cut2, npx_dyn + (threadIdx.x >> 6) * npx_dense_lds_floats<3>(),
                   [&](int j) { return XOW{XS[j], VT[j].w, OM[j].w}; },
                   [&](int qm) { return VortPair(S, qm, XS[qm], OM[qm].w); },
                   [&](int qm, const float* a) { commit_finish(S, qm, XS[qm], VX[qm], OM[qm], a, X, V); });
  }
}

// End of a step: every run member takes its head's committed state, density
// and lambda (the values the member would have computed, bit for bit).
__global__ void __launch_bounds__(NPX_BLOCK)
k_run_fill(const SceneC* __restrict__ scenes, const int* __restrict__ blk_scene, const Skin K,
           float4* __restrict__ X, float4* __restrict__ V, float* __restrict__ RHO, float* __restrict__ LAM) {
  const int s = blk_scene[npx_slot() / NPX_BLOCK];
  if (s < 0) return;
  const SceneC& S = scenes[s];
  const int q = npx_slot();
  if (q >= S.seg_off + S.n || !K.runs_s[s] || K.rstart[q]) return;
  const int h = K.ps[K.pi[q]];
  X[q] = X[h];
  V[q] = V[h];
  RHO[q] = RHO[h];
  LAM[q] = LAM[h];
}

// Grid diagnostics on x*_0: candidate count (duplicates kept) and in-range
// count (j != i, r^2 < h^2), plus the Teschner key of the particle's cell.
__global__ void __launch_bounds__(NPX_BLOCK)
k_debug_counts(const SceneC* __restrict__ scenes, const int* __restrict__ blk_scene, const int* __restrict__ clsmap,
               const int* __restrict__ cstart, const int* __restrict__ cell, const float4* __restrict__ XS,
               int* __restrict__ cand, int* __restrict__ nbr, unsigned long long* __restrict__ keys) {
  NPX_SCENE_PROLOGUE();
  const float4 xi = XS[q];
  int nc = 0, nn = 0;
  for_each_candidate(l, S.wx, wxy, cm, cstart, shift, [&](int j) {
    ++nc;
    if (j == q) return;
    const float4 xj = XS[j];
    if (sqn3(xi.x - xj.x, xi.y - xj.y, xi.z - xj.z) < S.h2) ++nn;
  });
  cand[q] = nc;
  nbr[q] = nn;
  const int lx = l % S.wx, t = l / S.wx, ly = t % S.wy, lz = t / S.wy;
  keys[q] = teschner_key(S.wlo0 + lx, S.wlo1 + ly, S.wlo2 + lz);
}

// Writes each particle's candidate list (as original indices) at out + off[q].
__global__ void __launch_bounds__(NPX_BLOCK)
k_debug_lists(const SceneC* __restrict__ scenes, const int* __restrict__ blk_scene, const int* __restrict__ clsmap,
              const int* __restrict__ cstart, const int* __restrict__ cell, const int* __restrict__ id,
              const long long* __restrict__ off, int* __restrict__ out) {
  NPX_SCENE_PROLOGUE();
  long long o = off[q];
  for_each_candidate(l, S.wx, wxy, cm, cstart, shift, [&](int j) { out[o++] = id[j]; });
}

// Per-scene health check: non-finite positions or velocities.
__global__ void __launch_bounds__(NPX_BLOCK)
k_scene_health(const SceneC* __restrict__ scenes, const int* __restrict__ blk_scene, const float4* __restrict__ X,
               const float4* __restrict__ V, int* __restrict__ flags) {
  const int s = blk_scene[npx_slot() / NPX_BLOCK];
  if (s < 0) return;
  const SceneC& S = scenes[s];
  const int q = npx_slot();
  if (q >= S.seg_off + S.n) return;
  const float4 x = X[q], v = V[q];
  const bool ok = isfinite(x.x) && isfinite(x.y) && isfinite(x.z) && isfinite(v.x) && isfinite(v.y) &&
                  isfinite(v.z);
  if (!ok) raise_flag(flags, s, NPX_FLAG_NONFINITE);
}

// ------------------------------------------------------ auxiliary grids
// Boundary particles and committed fluid positions get their own
// SpatialHashGrid in the reference (fluid_world.cpp:106-107, 134-135,
// boundary.cpp:162-163).  Points outside the window interior are left out
// (cls = -1): no point inside the window can have them within h.
// Offsets from bin_offset (blocks of NPX_BIN_BLOCK threads): the fluid's
// committed positions arrive in class order, and stacks of particles on one
// point would otherwise serialise on one counter.
__global__ void __launch_bounds__(NPX_BIN_BLOCK)
k_aux_bin(GridWin W, const float4* __restrict__ P, int n, const int* __restrict__ clsmap,
          int* __restrict__ ccount, int* __restrict__ cls_of, int* __restrict__ off_of) {
  __shared__ BinTable T;
  bin_clear(T);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  int c = -1;
  if (i < n) {
    const float4 p = P[i];
    const int lx = cell_coord(p.x, W.inv_h) - W.wlo0;
    const int ly = cell_coord(p.y, W.inv_h) - W.wlo1;
    const int lz = cell_coord(p.z, W.inv_h) - W.wlo2;
    if (!(lx < 1 || ly < 1 || lz < 1 || lx > W.wx - 2 || ly > W.wy - 2 || lz > W.wz - 2))
      c = clsmap[lx + W.wx * (ly + W.wy * lz)];
    cls_of[i] = c;
  }
  const bool active = c >= 0;
  const int off = bin_offset(T, c, active, ccount);
  if (active) off_of[i] = off;
}

// Provisional slot = class start + atomic offset; the point at offset 0 of
// a huge class registers it (hid, huge; its scene entry is -1) when hl < 31.
__global__ void k_aux_scatter(int n, const int* __restrict__ cstart, const int* __restrict__ cls_of,
                              const int* __restrict__ off_of, const int* __restrict__ key, int* __restrict__ tperm,
                              int* __restrict__ tkey, int* __restrict__ hid, int* __restrict__ huge, int hl) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int c = cls_of[i];
  if (c < 0) return;
  const int o = off_of[i];
  const int t = cstart[c] + o;
  tperm[t] = i;
  tkey[t] = key ? key[i] : i;
  if (o == 0 && is_huge(t, cstart[c + 1], hl)) {
    const int h = atomicAdd(huge, 1);
    hid[c] = h < NPX_HUGE_MAX ? h : -1;
    if (h < NPX_HUGE_MAX) {
      huge[1 + 3 * h] = -1;
      huge[2 + 3 * h] = t;
      huge[3 + 3 * h] = cstart[c + 1];
    }
  }
}

// Before an auxiliary sort with huge classes: no index is marked, no class
// registered.  hmark is cleared as nwords 32-bit words.
__global__ void k_huge_reset(unsigned* __restrict__ hmark_words, int nwords, int* __restrict__ huge) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i == 0) huge[0] = 0;
  if (i < nwords) hmark_words[i] = 0xFFFFFFFFu;
}

// Index-space rank of an auxiliary sort's huge classes: keys are unique in
// [0, nkeys) across the whole sort, one tile per 256 keys.
__global__ void __launch_bounds__(NPX_BLOCK)
k_aux_huge_hist(int nkeys, const int* __restrict__ huge, const unsigned char* __restrict__ hmark, int ntiles,
                int* __restrict__ hcnt) {
  const int nh = min(huge[0], NPX_HUGE_MAX);
  if (nh == 0) return;  // uniform
  const int ix = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  huge_hist_tile(nh, ix < nkeys, ix, hmark, ntiles, hcnt);
}

__global__ void __launch_bounds__(NPX_BLOCK)
k_aux_huge_place(int nkeys, const int* __restrict__ huge, const unsigned char* __restrict__ hmark,
                 const int* __restrict__ hidp, int ntiles, const int* __restrict__ hcnt, const float4* __restrict__ P,
                 int* __restrict__ perm, float4* __restrict__ PS) {
  const int nh = min(huge[0], NPX_HUGE_MAX);
  if (nh == 0) return;  // uniform
  const int ix = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  const int dst = huge_place_tile(nh, ix < nkeys, ix, huge, hmark, ntiles, hcnt);
  if (dst >= 0) {
    const int i = hidp[ix];
    perm[dst] = i;
    PS[dst] = P[i];
  }
}

// Sorts each 256-slot block of the binned slots that holds part of a class
// larger than NPX_RANK_DIRECT by (class, key), as k_rank_sort does for the
// fluid sort (keys are unique within a class).
__global__ void __launch_bounds__(NPX_BLOCK)
k_aux_rank_sort(int ncls, const int* __restrict__ cstart, const int* __restrict__ cls_of, int* __restrict__ tperm,
                int* __restrict__ tkey, unsigned char* __restrict__ bigb, const int* __restrict__ hid, int hl) {
  __shared__ unsigned long long sk[NPX_BLOCK];
  __shared__ int sp[NPX_BLOCK];
  const int i = (int)threadIdx.x;
  const int t = (int)blockIdx.x * NPX_BLOCK + i;
  __shared__ int slv;
  const bool active = t < cstart[ncls];
  if (i == 0) slv = 0;
  int p = 0, c = 0, lv = 0;
  bool large = false;
  if (active) {
    p = tperm[t];
    c = cls_of[p];
    const int b = cstart[c], e = cstart[c + 1];
    large = e - b > NPX_RANK_DIRECT && !(is_huge(b, e, hl) && hid[c] >= 0);  // ranked in index space
    if (large) lv = merge_levels(b, e);
  }
  const bool big = __syncthreads_or(large) != 0;
  if (lv > 0) atomicMax(&slv, lv);
  __syncthreads();
  if (i == 0) bigb[blockIdx.x] = (unsigned char)slv;  // read by the merge launches
  if (!big) return;
  unsigned long long key = ~0ull;
  if (active) key = ((unsigned long long)(unsigned)c << 32) | (unsigned)tkey[t];
  sk[i] = key;
  sp[i] = p;
  __syncthreads();
  for (int size = 2; size <= NPX_BLOCK; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      const int j = i ^ stride;
      if (j > i) {
        const unsigned long long a = sk[i], b = sk[j];
        if ((a > b) == ((i & size) == 0)) {
          sk[i] = b;
          sk[j] = a;
          const int tp = sp[i];
          sp[i] = sp[j];
          sp[j] = tp;
        }
      }
      __syncthreads();
    }
  }
  if (active) {
    tkey[t] = (int)(unsigned)sk[i];
    tperm[t] = sp[i];
  }
}

// Rank inside the class by key, as k_rank does for the fluid sort: small
// classes count directly, a larger class inside one block is already sorted,
// and one spanning several blocks is finished by k_aux_rank_merge.  Writes
// the sorted payload copy and each sorted slot's class (tcls).  256-thread
// blocks.
__global__ void __launch_bounds__(NPX_BLOCK)
k_aux_rank(int n, int ncls, const int* __restrict__ cstart, const int* __restrict__ cls_of,
           const int* __restrict__ tperm, const int* __restrict__ tkey, const float4* __restrict__ P,
           int* __restrict__ perm, float4* __restrict__ PS, int* __restrict__ tcls, const int* __restrict__ hid,
           int hl, unsigned char* __restrict__ hmark, int* __restrict__ hidp) {
  const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (t >= cstart[ncls]) return;
  const int i = tperm[t];
  const int c = cls_of[i];
  tcls[t] = c;
  const int b = cstart[c], e = cstart[c + 1];
  const int h = is_huge(b, e, hl) ? hid[c] : -1;
  if (h >= 0) {  // placed by k_aux_huge_place (the marks were cleared by k_huge_reset)
    hmark[tkey[t]] = (unsigned char)h;
    hidp[tkey[t]] = i;
  } else if (e - b <= NPX_RANK_DIRECT) {
    const int my = tkey[t];
    int r = 0;
    for (int u = b; u < e; ++u) r += (tkey[u] < my) ? 1 : 0;
    perm[b + r] = i;
    PS[b + r] = P[i];
  } else if (b / NPX_BLOCK == (e - 1) / NPX_BLOCK) {
    perm[t] = i;
    PS[t] = P[i];
  }
}

// Merge launch of the auxiliary sort's large classes (k_rank_merge without
// scenes); the class's last launch writes the permutation and the payload.
__global__ void __launch_bounds__(NPX_BLOCK)
k_aux_rank_merge(int ncls, const int* __restrict__ cstart, const int* __restrict__ tcls,
                 const unsigned char* __restrict__ bigb, int level, const int* __restrict__ pin,
                 const int* __restrict__ kin, int* __restrict__ pout, int* __restrict__ kout,
                 const float4* __restrict__ P, int* __restrict__ perm, float4* __restrict__ PS,
                 const int* __restrict__ hid, int hl) {
  if (bigb[blockIdx.x] <= level) return;  // uniform per block
  const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (t >= cstart[ncls]) return;
  const int c = tcls[t];
  const int b = cstart[c], e = cstart[c + 1];
  if (e - b <= NPX_RANK_DIRECT) return;
  const int npieces = (e - 1) / NPX_BLOCK - b / NPX_BLOCK + 1;
  if (npieces <= (1 << level)) return;
  if (is_huge(b, e, hl) && hid[c] >= 0) return;  // ranked in index space
  const int p = pin[t];
  const int my = kin[t];
  const int dst = merge_dst(b, e, t, level, my, kin);
  if (npieces <= (NPX_MERGE_FAN << level)) {
    perm[dst] = p;
    PS[dst] = P[p];
  } else {
    pout[dst] = p;
    kout[dst] = my;
  }
}

// ------------------------------------------------------ boundary sampling
// One shape as prepared on the host for k_sample_boundaries: its first output
// slot, its sample count and, per type, the loop bounds and constants the
// reference's samplers compute (boundary.cpp:15-127), with the same float
// operations.
struct ShapeS {
  int type, body, first, count;
  int n1[3], n2[3], axis_first[3];  // box: per face axis, grid dims and first local index
  int n;                            // sphere: point count; plane: points per side
  float he[3], r, golden_angle, extent;
  float nrm[3], offset, t1[3], t2[3];
  float pos[3], rot[4];             // shape local transform (rot: w, x, y, z)
};

__device__ __forceinline__ void cross3(const float* a, const float* b, float* o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}

// Transform::transform_point: position + q v q*, as v + w (2 u x v) + u x (2 u x v).
__device__ __forceinline__ void transform_point3(const float* pos, const float* rot, const float* v, float* o) {
  const float u[3] = {rot[1], rot[2], rot[3]};
  float uv[3], c[3];
  cross3(u, v, uv);
  for (int k = 0; k < 3; ++k) uv[k] = uv[k] + uv[k];
  cross3(u, uv, c);
  for (int k = 0; k < 3; ++k) o[k] = pos[k] + ((v[k] + rot[0] * uv[k]) + c[k]);
}

// One thread per boundary particle.  trig holds, per sphere sample, (sin
// theta, cos theta, sin phi, cos phi) evaluated on the host with the float
// functions of the host's libm, which the reference's sampler calls, so the
// samples are the host samplers' bit for bit.
__global__ void k_sample_boundaries(const ShapeS* __restrict__ shapes, int nshapes, int total, float spacing,
                                    const float4* __restrict__ trig, float* __restrict__ local3,
                                    int* __restrict__ body) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= total) return;
  int s = 0;
  while (s + 1 < nshapes && shapes[s + 1].first <= p) ++s;
  const ShapeS& S = shapes[s];
  const int i = p - S.first;
  float lp[3], o[3];
  if (S.type == NPX_SHAPE_BOX) {  // sample_box_boundary, boundary.cpp:15-53
    int axis = 0;
    while (axis < 2 && i >= S.axis_first[axis + 1]) ++axis;
    const int a1 = (axis + 1) % 3, a2 = (axis + 2) % 3;
    const int li = i - S.axis_first[axis], face = S.n1[axis] * S.n2[axis];
    const int sign = (li / face == 0) ? -1 : 1;
    const int i1 = (li % face) / S.n2[axis], i2 = (li % face) % S.n2[axis];
    lp[axis] = (float)sign * S.he[axis];
    lp[a1] = -S.he[a1] + (float)i1 * spacing;
    lp[a2] = -S.he[a2] + (float)i2 * spacing;
    transform_point3(S.pos, S.rot, lp, o);
  } else if (S.type == NPX_SHAPE_SPHERE) {  // sample_sphere_boundary, boundary.cpp:55-91
    const float4 tg = trig[p];
    const float st = tg.x, ct = tg.y, sp = tg.z, cp = tg.w;
    lp[0] = S.r * st * cp;
    lp[1] = S.r * st * sp;
    lp[2] = S.r * ct;
    transform_point3(S.pos, S.rot, lp, o);
  } else {  // sample_plane_boundary, boundary.cpp:93-126 (world-owned: no transform)
    const float a = -S.extent + (float)(i / S.n) * spacing, c = -S.extent + (float)(i % S.n) * spacing;
    for (int k = 0; k < 3; ++k) o[k] = (S.nrm[k] * S.offset + S.t1[k] * a) + S.t2[k] * c;
  }
  local3[3 * (size_t)p] = o[0];
  local3[3 * (size_t)p + 1] = o[1];
  local3[3 * (size_t)p + 2] = o[2];
  body[p] = S.body;
}

// Cell (local window index) of an arbitrary point, or -1 if its 27-cell
// neighbourhood leaves the window.
__device__ __forceinline__ int aux_cell(const GridWin& W, float x, float y, float z) {
  const int lx = cell_coord(x, W.inv_h) - W.wlo0;
  const int ly = cell_coord(y, W.inv_h) - W.wlo1;
  const int lz = cell_coord(z, W.inv_h) - W.wlo2;
  if (lx < 1 || ly < 1 || lz < 1 || lx > W.wx - 2 || ly > W.wy - 2 || lz > W.wz - 2) return -1;
  return lx + W.wx * (ly + W.wy * lz);
}

// compute_boundary_volumes (boundary.cpp:152-176): psi_b = 1 / sum W, over
// the boundary candidates of b (self and every body included).
__global__ void k_psi(GridWin W, float h2, float c6, const float4* __restrict__ BP, int nb,
                      const int* __restrict__ clsmap, const int* __restrict__ cstart, const float4* __restrict__ PS,
                      float* __restrict__ psi) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  const float4 xb = BP[b];
  const int l = aux_cell(W, xb.x, xb.y, xb.z);
  float sum = 0.0f;
  if (l >= 0) {
    for_each_candidate(l, W.wx, W.wx * W.wy, clsmap, cstart, 0, [&](int j) {
      const float4 xj = PS[j];
      sum += poly6(sqn3(xb.x - xj.x, xb.y - xj.y, xb.z - xj.z), h2, c6);
    });
  }
  psi[b] = (sum > 1e-10f) ? 1.0f / sum : 0.0f;
}

// boundary_world_positions (boundary.cpp:178-192): body pose applied with
// Eigen's quaternion*vector form; world-owned particles keep their position.
// Output w = psi.
__global__ void k_bpos(const float4* __restrict__ BL, const int* __restrict__ body, const float* __restrict__ psi,
                       int nb, const float* __restrict__ pose7, int nbodies, float4* __restrict__ BP) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  const float4 p = BL[b];
  const int bi = body[b];
  float ox = p.x, oy = p.y, oz = p.z;
  if (bi >= 0 && bi < nbodies) {
    const float* P = pose7 + 7 * bi;
    const float qw = P[3], qx = P[4], qy = P[5], qz = P[6];
    float ux = qy * p.z - qz * p.y, uy = qz * p.x - qx * p.z, uz = qx * p.y - qy * p.x;
    ux = ux + ux;
    uy = uy + uy;
    uz = uz + uz;
    const float cx = qy * uz - qz * uy, cy = qz * ux - qx * uz, cz = qx * uy - qy * ux;
    const float rx = (p.x + qw * ux) + cx, ry = (p.y + qw * uy) + cy, rz = (p.z + qw * uz) + cz;
    ox = P[0] + rx;
    oy = P[1] + ry;
    oz = P[2] + rz;
  }
  BP[b] = make_float4(ox, oy, oz, psi[b]);
}

// k_bpos over the batch's boundary pool: bscene[b] < 0 marks a free
// position (its output is never read); a particle's body is its scene's first
// body (sbase) + its local body.
__global__ void k_bpos_b(const float4* __restrict__ BL, const int* __restrict__ lbody, const int* __restrict__ bscene,
                         const int* __restrict__ sbase, const float* __restrict__ psi, int nb,
                         const float* __restrict__ pose7, int nbodies, float4* __restrict__ BP) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  const int s = bscene[b];
  if (s < 0) return;
  const float4 p = BL[b];
  const int lb = lbody[b];
  const int bi = (lb >= 0) ? sbase[s] + lb : -1;
  float ox = p.x, oy = p.y, oz = p.z;
  if (bi >= 0 && bi < nbodies) {
    const float* P = pose7 + 7 * bi;
    const float qw = P[3], qx = P[4], qy = P[5], qz = P[6];
    float ux = qy * p.z - qz * p.y, uy = qz * p.x - qx * p.z, uz = qx * p.y - qy * p.x;
    ux = ux + ux;
    uy = uy + uy;
    uz = uz + uz;
    const float cx = qy * uz - qz * uy, cy = qz * ux - qx * uz, cz = qx * uy - qy * ux;
    const float rx = (p.x + qw * ux) + cx, ry = (p.y + qw * uy) + cy, rz = (p.z + qw * uz) + cz;
    ox = P[0] + rx;
    oy = P[1] + ry;
    oz = P[2] + rz;
  }
  BP[b] = make_float4(ox, oy, oz, psi[b]);
}

// FluidWorld::apply_boundary_density (fluid_world.cpp:101-123) on the
// committed positions: rho_i += sum_b (rho0 psi_b) W(x_i - x_b).
__global__ void __launch_bounds__(NPX_BLOCK)
k_boundary_density(const SceneC* __restrict__ scenes, const int* __restrict__ blk_scene, GridWin W,
                   const float4* __restrict__ X, const int* __restrict__ bclsmap, const int* __restrict__ bcstart,
                   const float4* __restrict__ BPS, float* __restrict__ RHO, int* __restrict__ flags) {
  const int s = blk_scene[npx_slot() / NPX_BLOCK];
  if (s < 0) return;
  const SceneC& S = scenes[s];
  const int q = npx_slot();
  if (q >= S.seg_off + S.n) return;
  const float4 xi = X[q];
  const int l = aux_cell(W, xi.x, xi.y, xi.z);
  if (l < 0) {
    raise_flag(flags, s, NPX_FLAG_WINDOW);
    return;
  }
  float bd = 0.0f;
  const float h2 = S.h2, c6 = S.c6, rho0 = S.rho0;
  for_each_candidate(l, W.wx, W.wx * W.wy, bclsmap, bcstart, 0, [&](int j) {
    const float4 xb = BPS[j];
    bd += rho0 * xb.w * poly6(sqn3(xi.x - xb.x, xi.y - xb.y, xi.z - xb.z), h2, c6);
  });
  RHO[q] = RHO[q] + bd;
}

// Payload for the committed-fluid grid: (x, rho) with key = original index.
__global__ void __launch_bounds__(NPX_BLOCK)
k_fluid_payload(const SceneC* __restrict__ scenes, const int* __restrict__ blk_scene, const float4* __restrict__ X,
                const float* __restrict__ RHO, float4* __restrict__ P, const int* __restrict__ id,
                int* __restrict__ gkey) {
  const int s = blk_scene[npx_slot() / NPX_BLOCK];
  if (s < 0) return;
  const SceneC& S = scenes[s];
  const int q = npx_slot();
  if (q >= S.seg_off + S.n) return;
  const float4 x = X[q];
  P[q] = make_float4(x.x, x.y, x.z, RHO[q]);
  gkey[q] = S.seg_off + id[q];  // unique over the batch, ordered like the index within a scene
}

// Runs of identical points in the coupling's fluid sort: slot j starts a run
// unless it holds the same class and bitwise the same payload (committed
// position and density) as slot j - 1.  F[j] = 1 at a start, 0 for the
// slots past the sorted count cstart[ncls] and for F[n].
__global__ void k_payload_runs(int n, const int* __restrict__ cstart, int ncls, const int* __restrict__ tcls,
                               const float4* __restrict__ PS, int* __restrict__ F) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j > n) return;
  int f = 0;
  if (j < cstart[ncls]) {
    f = 1;
    if (j > 0 && tcls[j] == tcls[j - 1]) {
      const float4 a = PS[j], b = PS[j - 1];
      f = (__float_as_uint(a.x) != __float_as_uint(b.x) || __float_as_uint(a.y) != __float_as_uint(b.y) ||
           __float_as_uint(a.z) != __float_as_uint(b.z) || __float_as_uint(a.w) != __float_as_uint(b.w))
              ? 1
              : 0;
    }
  }
  F[j] = f;
}

// After the exclusive scan of F: pi[j] = the run of slot j, ps[r] = the
// first slot of run r.
__global__ void k_payload_pieces(int n, const int* __restrict__ cstart, int ncls, const int* __restrict__ E,
                                 int* __restrict__ pi, int* __restrict__ ps) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n || j >= cstart[ncls]) return;
  pi[j] = E[j + 1] - 1;
  if (E[j + 1] != E[j]) ps[E[j]] = j;
}

// The coupling force term of fluid point xf on boundary particle xb
// (fluid_world.cpp:140-160): false when the pair adds nothing (beyond h, or
// outside the kernel's support).
__device__ __forceinline__ bool coupling_term(float4 xb, float4 xf, float h, float h2, float cs, float rho0,
                                              float mass, float& tx, float& ty, float& tz) {
  const float psi_b = xb.w;
  const float rx = xb.x - xf.x, ry = xb.y - xf.y, rz = xb.z - xf.z;
  const float r_sq = sqn3(rx, ry, rz);
  if (r_sq >= h2) return false;
  const float rho_f = (xf.w < 1.0f) ? 1.0f : xf.w;
  const float pr = rho_f - rho0;
  const float pressure = (0.0f < pr) ? pr : 0.0f;
  const float pc = mass * pressure / (rho_f * rho_f);
  bool in;
  const float co = spiky_coeff(r_sq, h, cs, &in);
  if (!in) return false;
  const float kc = pc * rho0 * psi_b;
  tx = kc * (co * rx);
  ty = kc * (co * ry);
  tz = kc * (co * rz);
  return true;
}

// F_b of one boundary particle in window cell l, computed by the calling
// wave (every lane gets the sum).  The fluid candidates of b's 27 cells are
// taken in the reference's order as runs of identical points (rpi / rps:
// k_payload_pieces; a candidate range's first slot always starts a run), 64
// runs at a time: each lane forms one run's term; the in-range terms go to
// LDS compacted in run order (mbcnt of the ballot), and lanes 0 / 1 / 2 add
// the x / y / z row in order, a run of k as k additions in a row (add_k), so
// each component is the reference's one-by-one sum bit for bit.  The three
// sums advance side by side at one LDS read and one addition per term, which
// matters next to piles of fluid: a boundary particle there has thousands of
// in-range terms in one serial sum.  Callers run 256-thread blocks (at most
// four waves; each wave uses its own LDS rows).
__device__ __forceinline__ void coupling_wave(float4 xb, int l, int wx, int wxy, const int* __restrict__ clsmap,
                                              const int* __restrict__ cstart, const int* __restrict__ rpi,
                                              const int* __restrict__ rps, const float4* __restrict__ FS, float h,
                                              float h2, float cs, float rho0, float mass, float& fx, float& fy,
                                              float& fz) {
  __shared__ float cw_t[4][3][64];  // per wave: the chunk's in-range terms, one row per component
  __shared__ int cw_k[4][64];       // per wave: their run lengths
  const int w = (int)(threadIdx.x >> 6);
  float* __restrict__ rt = &cw_t[w][0][0];
  int* __restrict__ rk = &cw_k[w][0];
  const int lane = (int)(threadIdx.x & 63);
  float acc = 0.0f;  // lane u < 3: component u of F_b
  auto range = [&](int bgn, int e) {
    if (bgn >= e) return;
    const int r0 = rpi[bgn], r1 = rpi[e - 1] + 1;
    for (int r = r0; r < r1; r += 64) {
      const int ri = r + lane;
      float tx = 0.0f, ty = 0.0f, tz = 0.0f;
      int k = 0;
      bool has = false;
      if (ri < r1) {
        const int j = rps[ri];
        k = ((ri + 1 < r1) ? rps[ri + 1] : e) - j;
        has = coupling_term(xb, FS[j], h, h2, cs, rho0, mass, tx, ty, tz);
      }
      const unsigned long long m = __ballot(has);
      if (has) {
        const int pos = (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
        rt[pos] = tx;
        rt[64 + pos] = ty;
        rt[128 + pos] = tz;
        rk[pos] = k;
      }
      npx_wave_lds_sync();
      const int n = (int)__builtin_popcountll(m);
      if (lane < 3) {
        const float* __restrict__ row = rt + 64 * lane;
        for (int i = 0; i < n; ++i) add_k(acc, row[i], rk[i]);
      }
      npx_wave_lds_sync();
    }
  };
#pragma unroll 1
  for (int dz = -1; dz <= 1; ++dz) {
#pragma unroll 1
    for (int dy = -1; dy <= 1; ++dy) {
      const int base = l + dy * wx + dz * wxy;
      const int c0 = clsmap[base - 1];
      const int c1 = clsmap[base];
      const int c2 = clsmap[base + 1];
      if (c1 == c0 + 1 && c2 == c0 + 2) {
        range(cstart[c0], cstart[c0 + 3]);
      } else {
        range(cstart[c0], cstart[c0 + 1]);
        range(cstart[c1], cstart[c1 + 1]);
        range(cstart[c2], cstart[c2 + 1]);
      }
    }
  }
  fx = __shfl(acc, 0, 64);
  fy = __shfl(acc, 1, 64);
  fz = __shfl(acc, 2, 64);
}

// FluidWorld::apply_coupling_forces, per boundary particle
// (fluid_world.cpp:140-160): F_b = sum_f ((m p_f / rho_f^2) rho0 psi_b) gradW,
// one wave per boundary particle (coupling_wave).
__global__ void __launch_bounds__(256)
k_coupling_particles(GridWin W, float h, float h2, float cs, float rho0, float mass, const float4* __restrict__ BP,
                     const int* __restrict__ body, int nb, const uint8_t* __restrict__ is_static, int nbodies,
                     const int* __restrict__ fclsmap, const int* __restrict__ fcstart,
                     const float4* __restrict__ FS, const int* __restrict__ rpi, const int* __restrict__ rps,
                     float4* __restrict__ FB) {
  const int b = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  if (b >= nb) return;  // uniform per wave
  const int bi = body[b];
  float fx = 0.0f, fy = 0.0f, fz = 0.0f;
  if (bi >= 0 && bi < nbodies && !is_static[bi]) {
    const float4 xb = BP[b];
    const int l = aux_cell(W, xb.x, xb.y, xb.z);
    if (l >= 0)
      coupling_wave(xb, l, W.wx, W.wx * W.wy, fclsmap, fcstart, rpi, rps, FS, h, h2, cs, rho0, mass, fx, fy, fz);
  }
  if ((threadIdx.x & 63) == 0) FB[b] = make_float4(fx, fy, fz, 0.0f);
}

// Per-body sums in boundary-particle order (World::apply_force / apply_torque,
// world.cpp:189-216): one 64-lane block per body fetches 64 of its particles
// at a time, each lane forming its particle's force and torque terms, and the
// terms are added in ascending particle order (broadcast lane by lane), so
// the float accumulation order is the reference's and the result is
// deterministic.  Body k's particles are body_idx[bstart[k] .. bend[k]).
__global__ void __launch_bounds__(64)
k_body_reduce(const int* __restrict__ bstart, const int* __restrict__ bend, const int* __restrict__ body_idx,
              int nbodies,
              const float4* __restrict__ FB, const float4* __restrict__ BP, const float* __restrict__ pose7,
              float* __restrict__ out6) {
  const int k = (int)blockIdx.x;
  if (k >= nbodies) return;  // uniform per block
  const int t = (int)threadIdx.x;
  const float cxm = pose7[7 * k], cym = pose7[7 * k + 1], czm = pose7[7 * k + 2];
  float Fx = 0.0f, Fy = 0.0f, Fz = 0.0f, Tx = 0.0f, Ty = 0.0f, Tz = 0.0f;
  const int end = bend[k];
  for (int u0 = bstart[k]; u0 < end; u0 += 64) {
    const int u = u0 + t;
    float fx = 0.0f, fy = 0.0f, fz = 0.0f, tx = 0.0f, ty = 0.0f, tz = 0.0f;
    bool use = false;
    if (u < end) {
      const int b = body_idx[u];
      const float4 f = FB[b];
      use = sqn3(f.x, f.y, f.z) > 1e-20f;
      if (use) {
        const float4 xb = BP[b];
        const float ax = xb.x - cxm, ay = xb.y - cym, az = xb.z - czm;
        fx = f.x;
        fy = f.y;
        fz = f.z;
        tx = ay * f.z - az * f.y;
        ty = az * f.x - ax * f.z;
        tz = ax * f.y - ay * f.x;
      }
    }
    for (unsigned long long m = __ballot(use); m; m &= m - 1ull) {
      const int l = (int)__builtin_ctzll(m);
      Fx += __shfl(fx, l, 64);
      Fy += __shfl(fy, l, 64);
      Fz += __shfl(fz, l, 64);
      Tx += __shfl(tx, l, 64);
      Ty += __shfl(ty, l, 64);
      Tz += __shfl(tz, l, 64);
    }
  }
  if (t == 0) {
    out6[6 * k + 0] = Fx;
    out6[6 * k + 1] = Fy;
    out6[6 * k + 2] = Fz;
    out6[6 * k + 3] = Tx;
    out6[6 * k + 4] = Ty;
    out6[6 * k + 5] = Tz;
  }
}

// ------------------------------------------------- batched Akinci coupling
// The boundary particles of every scene of a batch context are binned into
// their own scene's cell window and alias classes, the same buckets the
// single-scene path uses (its auxiliary grids also take the scene's window),
// so boundary density, coupling forces and their float order are the
// single-scene ones.

// Window-local cell of a point in scene S's window, or -1 when its 27-cell
// neighbourhood leaves the window.
__device__ __forceinline__ int scene_cell(const SceneC& S, float x, float y, float z) {
  const int lx = cell_coord(x, S.inv_h) - S.wlo0;
  const int ly = cell_coord(y, S.inv_h) - S.wlo1;
  const int lz = cell_coord(z, S.inv_h) - S.wlo2;
  if (lx < 1 || ly < 1 || lz < 1 || lx > S.wx - 2 || ly > S.wy - 2 || lz > S.wz - 2) return -1;
  return lx + S.wx * (ly + S.wy * lz);
}

// Bins points of known scene (pscene[i], -1: none) into global classes.
__global__ void __launch_bounds__(NPX_BIN_BLOCK)
k_bin_scene_points(const SceneC* __restrict__ scenes, const float4* __restrict__ P, const int* __restrict__ pscene,
                   int n, const int* __restrict__ clsmap, int* __restrict__ ccount, int* __restrict__ cls_of,
                   int* __restrict__ off_of) {
  __shared__ BinTable T;
  bin_clear(T);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  int c = -1;
  if (i < n) {
    const int s = pscene[i];
    if (s >= 0) {
      const SceneC& S = scenes[s];
      const float4 p = P[i];
      const int l = scene_cell(S, p.x, p.y, p.z);
      if (l >= 0) c = clsmap[S.cls_base + l];
    }
    cls_of[i] = c;
  }
  const int off = bin_offset(T, c, c >= 0, ccount);
  if (c >= 0) off_of[i] = off;
}

// Scene of each fluid slot when its scene has a boundary, else -1.
__global__ void __launch_bounds__(NPX_BLOCK)
k_coupled_slots(const SceneC* __restrict__ scenes, const int* __restrict__ blk_scene, int* __restrict__ pscene) {
  const int q = npx_slot();
  const int s = blk_scene[q / NPX_BLOCK];
  int r = -1;
  if (s >= 0) {
    const SceneC& S = scenes[s];
    if (q < S.seg_off + S.n && S.nbnd > 0) r = s;
  }
  pscene[q] = r;
}

// FluidWorld::apply_boundary_density (fluid_world.cpp:101-123) for every
// scene with a boundary: rho_i += sum_b (rho0 psi_b) W(x_i - x_b), committed x.
__global__ void __launch_bounds__(NPX_BLOCK)
k_boundary_density_b(const SceneC* __restrict__ scenes, const int* __restrict__ blk_scene,
                     const float4* __restrict__ X, const int* __restrict__ clsmap, const int* __restrict__ bcstart,
                     const float4* __restrict__ BPS, float* __restrict__ RHO, int* __restrict__ flags) {
  const int s = blk_scene[npx_slot() / NPX_BLOCK];
  if (s < 0) return;
  const SceneC& S = scenes[s];
  const int q = npx_slot();
  if (q >= S.seg_off + S.n || S.nbnd == 0) return;
  const float4 xi = X[q];
  const int l = scene_cell(S, xi.x, xi.y, xi.z);
  if (l < 0) {
    raise_flag(flags, s, NPX_FLAG_WINDOW);
    return;
  }
  float bd = 0.0f;
  const float h2 = S.h2, c6 = S.c6, rho0 = S.rho0;
  for_each_candidate(l, S.wx, S.wx * S.wy, clsmap + S.cls_base, bcstart, 0, [&](int j) {
    const float4 xb = BPS[j];
    bd += rho0 * xb.w * poly6(sqn3(xi.x - xb.x, xi.y - xb.y, xi.z - xb.z), h2, c6);
  });
  RHO[q] = RHO[q] + bd;
}

// FluidWorld::apply_coupling_forces per boundary particle
// (fluid_world.cpp:140-160), every scene: F_b = sum_f ((m p_f / rho_f^2)
// rho0 psi_b) gradW over the committed-fluid buckets of b's scene, one wave
// per boundary particle (coupling_wave).
__global__ void __launch_bounds__(256)
k_coupling_b(const SceneC* __restrict__ scenes, const float4* __restrict__ BP, const int* __restrict__ bscene,
             const int* __restrict__ lbody, const int* __restrict__ sbase, int nb,
             const uint8_t* __restrict__ is_static, int nbodies,
             const int* __restrict__ clsmap, const int* __restrict__ fcstart, const float4* __restrict__ FS,
             const int* __restrict__ rpi, const int* __restrict__ rps, float4* __restrict__ FB) {
  const int b = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  if (b >= nb) return;  // uniform per wave
  const int s = bscene[b];
  if (s < 0) return;  // a free pool position
  const int lb = lbody[b];
  const int bi = (lb >= 0) ? sbase[s] + lb : -1;
  float fx = 0.0f, fy = 0.0f, fz = 0.0f;
  if (bi >= 0 && bi < nbodies && !is_static[bi]) {
    const SceneC& S = scenes[s];
    const float4 xb = BP[b];
    const int l = scene_cell(S, xb.x, xb.y, xb.z);
    if (l >= 0)
      coupling_wave(xb, l, S.wx, S.wx * S.wy, clsmap + S.cls_base, fcstart, rpi, rps, FS, S.h, S.h2, S.cs, S.rho0,
                    S.mass, fx, fy, fz);
  }
  if ((threadIdx.x & 63) == 0) FB[b] = make_float4(fx, fy, fz, 0.0f);
}

// AoS float triples -> float4 (w = 0), for points copied to the device raw.
__global__ void k_pack_float4(const float* __restrict__ p3, int n, float4* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = make_float4(p3[3 * (size_t)i], p3[3 * (size_t)i + 1], p3[3 * (size_t)i + 2], 0.0f);
}

// Widens int32 offsets to the int64 offsets the list writer takes.
__global__ void k_widen_i32(const int* __restrict__ in, int n, long long* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (long long)in[i];
}

// query_neighbors for arbitrary points (standalone SpatialHashGrid API).
__global__ void k_query_count(GridWin W, const float4* __restrict__ Q, int nq, const int* __restrict__ clsmap,
                              const int* __restrict__ cstart, int* __restrict__ counts) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nq) return;
  const float4 p = Q[i];
  const int l = aux_cell(W, p.x, p.y, p.z);
  int c = 0;
  if (l >= 0) for_each_candidate(l, W.wx, W.wx * W.wy, clsmap, cstart, 0, [&](int) { ++c; });
  counts[i] = c;
}

__global__ void k_query_lists(GridWin W, const float4* __restrict__ Q, int nq, const int* __restrict__ clsmap,
                              const int* __restrict__ cstart, const int* __restrict__ perm,
                              const long long* __restrict__ off, int* __restrict__ out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nq) return;
  const float4 p = Q[i];
  const int l = aux_cell(W, p.x, p.y, p.z);
  long long o = off[i];
  if (l >= 0) for_each_candidate(l, W.wx, W.wx * W.wy, clsmap, cstart, 0, [&](int j) { out[o++] = perm[j]; });
}
