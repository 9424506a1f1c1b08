// npx_engine.hpp -- host side of libnpx_hip: device buffers, scene table,
// launch sequence of one PBF step, Akinci coupling and the C ABI.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/novaphy_hip.h"
#include "npx_kernels.hpp"

namespace npx {

static const float kPI = 3.14159265358979323846f;  // math_utils.h:53

// HIP error -> C ABI code: device memory exhaustion is NPX_ENOMEM, any other
// runtime error NPX_EDEVICE.
#define NPX_HIP_CODE(e) ((e) == hipErrorOutOfMemory ? NPX_ENOMEM : NPX_EDEVICE)

#define NPX_HIP(expr)                                                                        \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess) return fail(NPX_HIP_CODE(e_), "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                                      __FILE__, __LINE__);                                   \
  } while (0)

#define NPX_HIP_C(c, expr)                                                                   \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess)                                                                    \
      return (c)->fail(NPX_HIP_CODE(e_), "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
  } while (0)

// Largest particle-slot capacity of a context: slot indices, list words and
// block counts stay inside int32 arithmetic (2^28 slots need about 200 GB of
// state and lists, within one MI355X's 288 GB).
#define NPX_MAX_SLOTS (1 << 28)

#define NPX_TRY(expr)        \
  do {                       \
    int r_ = (expr);         \
    if (r_ != NPX_OK) return r_; \
  } while (0)

template <class T>
struct DBuf {  // plain device buffer
  T* p = nullptr;
  size_t n = 0;
  hipError_t ensure(size_t need) {
    if (need <= n) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
    hipError_t e = hipMalloc((void**)&p, std::max<size_t>(need, 1) * sizeof(T));
    if (e == hipSuccess) n = need;
    return e;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
};

// First-fit allocator over [0, cap) in abstract units.
struct RangeAlloc {
  int cap = 0;
  std::map<int, int> free_;  // start -> length
  void init(int c) {
    cap = c;
    free_.clear();
    if (c > 0) free_[0] = c;
  }
  int alloc(int len) {
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= len) {
        int s = it->first, l = it->second;
        free_.erase(it);
        if (l > len) free_[s + len] = l - len;
        return s;
      }
    }
    return -1;
  }
  void release(int s, int len) {
    auto it = free_.emplace(s, len).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
};

struct SceneH {
  bool used = false;
  int nbnd = 0;  // boundary particles (batched coupling)
  npx_pbf_params prm{};
  float dt = 0.0f, g[3] = {0, 0, 0}, mass = 0.0f;
  int n = 0;
  int seg_off = 0, seg_blocks = 0;
  int cls_base = 0, ncell = 0;
  int wlo[3] = {0, 0, 0}, wdim[3] = {0, 0, 0};
  bool window_valid = false;
};

// Smallest float x with sqrtf(x) >= h (so for r^2 >= max(h^2, x) both poly6
// and spiky_grad of the reference return zero).
static float spiky_cut(float h) {
  float x = h * h;
  // walk down while sqrt still reaches h, then up until it does
  while (x > 0.0f && std::sqrt(std::nextafter(x, 0.0f)) >= h) x = std::nextafter(x, 0.0f);
  while (std::sqrt(x) < h) x = std::nextafter(x, 1e30f);
  return x;
}

// Largest float x >= 0 with sqrtf(x) <= c.  sqrtf is correctly rounded on the
// host and in the kernels (monotone), so sqrtf(r^2) <= c  <=>  r^2 <= x.
static float sqrt_le_max(float c) {
  float x = c * c;
  while (std::sqrt(std::nextafter(x, 1e30f)) <= c) x = std::nextafter(x, 1e30f);
  while (x > 0.0f && std::sqrt(x) > c) x = std::nextafter(x, 0.0f);
  return x;
}

static SceneC make_scene_consts(const SceneH& H, float skin_frac) {
  const npx_pbf_params& P = H.prm;
  SceneC S;
  std::memset(&S, 0, sizeof(S));
  const float h = P.kernel_radius;
  S.h = h;
  S.h2 = h * h;
  S.inv_h = 1.0f / h;
  const float h2 = h * h;
  const float h9 = h2 * h2 * h2 * h2 * h;
  const float k64pi = 64.0f * kPI;
  S.c6 = 315.0f / (k64pi * h9);
  const float h6 = h * h * h * h * h * h;
  S.cs = 45.0f / (kPI * h6);
  S.rho0 = P.rest_density;
  S.inv_rho0 = 1.0f / P.rest_density;
  S.mass = H.mass;
  S.m_inv_rho0 = H.mass * S.inv_rho0;
  S.eps = P.epsilon;
  S.xsph_c = P.xsph_viscosity;
  S.xsph_on = P.xsph_viscosity > 0.0f ? 1 : 0;
  S.vort_eps = P.vorticity_epsilon;
  S.vort_on = P.vorticity_epsilon != 0.0f ? 1 : 0;
  S.corr_k = P.corr_k;
  S.corr_n = P.corr_n;
  {
    const float dq = P.corr_delta_q * h;
    const float r_sq = dq * dq;
    float w = 0.0f;
    if (!(r_sq >= h * h)) {
      const float diff = h2 - r_sq;
      w = S.c6 * diff * diff * diff;
    }
    S.wdq = w;
  }
  S.dt = H.dt;
  S.inv_dt = 1.0f / H.dt;
  S.max_speed = h / H.dt;
  S.max_speed_sq = S.max_speed * S.max_speed;
  S.gdt0 = H.g[0] * H.dt;
  S.gdt1 = H.g[1] * H.dt;
  S.gdt2 = H.g[2] * H.dt;
  const float pr = P.particle_radius;
  S.clo0 = P.domain_lower[0] + pr;
  S.clo1 = P.domain_lower[1] + pr;
  S.clo2 = P.domain_lower[2] + pr;
  S.chi0 = P.domain_upper[0] - pr;
  S.chi1 = P.domain_upper[1] - pr;
  S.chi2 = P.domain_upper[2] - pr;
  S.sp_hi = spiky_cut(h);
  S.sp_lo = sqrt_le_max(1e-7f);
  S.fdiv = (h >= 0.015625f && h <= 64.0f) ? 1 : 0;  // operand range of npx_div_in
  S.cut2 = std::max(S.h2, S.sp_hi);
  {
    // skin radius hs = (1 + f) h; a pair outside the skin at x*_0 stays
    // beyond the cut while c_i + c_j <= hs - |cut|, c a particle's summed
    // displacement since.  The relative margins cover the rounding of the
    // float distance tests and sums.
    const double hs = (double)h * (1.0 + (double)skin_frac);
    S.skin_cut2 = (float)(hs * hs);
    const double allow = hs * (1.0 - 1e-5) - std::sqrt((double)S.cut2) * (1.0 + 1e-5);
    S.skin_allow = allow > 0.0 ? (float)allow : -1.0f;
  }
  S.iters = P.solver_iterations;
  S.bounds = P.use_domain_bounds ? 1 : 0;
  S.wlo0 = H.wlo[0];
  S.wlo1 = H.wlo[1];
  S.wlo2 = H.wlo[2];
  S.wx = H.wdim[0];
  S.wy = H.wdim[1];
  S.wz = H.wdim[2];
  S.cls_base = H.cls_base;
  S.ncell = H.ncell;
  S.seg_off = H.seg_off;
  S.n = H.n;
  S.seg_blocks = H.seg_blocks;
  S.nbnd = H.nbnd;
  return S;
}

static inline int cdiv(long long a, long long b) { return (int)((a + b - 1) / b); }

struct KTimer {
  std::string name;
  hipEvent_t a, b;
};

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  uint32_t flags = 0;
  bool fast = false;
  bool batch = false;
  std::string err;

  int slot_cap = 0;   // slots (multiple of NPX_BLOCK)
  int cls_cap = 0;    // classes
  int max_scenes = 1;
  RangeAlloc slot_alloc;  // in blocks
  RangeAlloc cls_alloc;   // in cells
  std::vector<SceneH> scenes;
  int hw_blocks = 0;
  bool table_dirty = true;
  // tunables (results are bitwise independent of them)
  int nbr_block = 128;  // block size of the neighbour passes (64..NPX_MAX_BLOCK)
  int persist_blocks = 0;  // > 0: at most this many per-lane blocks per neighbour pass, each looping
  int huge_levels = NPX_HUGE_LEVELS;  // classes spanning more than 2^this blocks rank in index space
  int list_policy = 0;  // 0: no skin lists (every testing pass sweeps inline); 1: skin lists where they held
  int xcd_remap = 1;    // 1: XCD-aware block order in the neighbour passes
  float skin_frac = 0.1f;  // skin radius of the per-step candidate lists, as a fraction of h
  std::vector<int> h_blk_scene;

  // particle arrays (slot order)
  DBuf<float4> X, X2, V, V2, XS0, XS1, VT, OM, VX;
  DBuf<float> RHO, RHO2, LAM, LAM2;
  DBuf<int> ID, ID2, CELL, CELL2, CLS, OFF, TPERM, TKEY, PERM;
  DBuf<int> MPERM, MKEY;  // the other slot / key buffers of the large-class merge launches
  DBuf<int> CRF, CRPI, CRPS;  // runs of the coupling's fluid sort: flags / scan, run of each slot, run starts
  DBuf<uint8_t> BIGB;     // per 256-slot block: merge launches its large classes need
  // index-space rank of the huge classes: per original index (seg_off +
  // index) its huge class and its unsorted slot; per huge class and tile its
  // member count (then offset); the step's huge-class count, scenes and slot ranges
  DBuf<uint8_t> HMARK;
  DBuf<int> HIDP, HCNT, HUGE;
  DBuf<int> XORD;  // the step's XCD order of the neighbour passes (k_pieces)
  DBuf<int> SKL, SKN;  // per-step skin lists (NPX_SKIN_CAP per slot) and counts
  DBuf<int> EXL, EXN;  // in-range lists of the current testing pass and counts
  DBuf<float> CDB;  // summed displacement per slot since the skin lists were built
  DBuf<int> CMAX;   // per class: float bits of the largest summed displacement; then per class a run flag
  DBuf<uint8_t> RSTART;  // per slot: 1 if the slot starts a run of identical state
  DBuf<int> PI, PS;      // per slot: its run; per run: its first slot
  DBuf<int> BOFF;        // per 256-slot block: run starts, then their exclusive prefix
  DBuf<int> DL, DN;      // dense list (heads whose skin list overflowed) and its length
  DBuf<int> LST;         // per-scene skin-list statistics and build flags (Skin::lst)
  // class tables
  DBuf<int> clsmap, cstart, bsum;
  DBuf<unsigned long long> htk;
  DBuf<int> htv;
  DBuf<SceneC> d_scenes;
  DBuf<int> d_blk_scene;
  DBuf<int> d_flags;
  int* h_flags = nullptr;  // pinned
  int xs_final = 0;

  // Pinned staging of the per-step body transfers (the NB x 7 poses and
  // is_static flags in, the NB x 6 forces and torques out), so uploads are
  // asynchronous DMA on the stream and the host blocks once per coupling
  // call, at the readback.  ev_up marks the last upload from the staging
  // buffers; it is waited on before the host overwrites them.
  float* h_pose = nullptr;
  uint8_t* h_static = nullptr;
  float* h_out = nullptr;
  int h_body_cap = 0;
  hipEvent_t ev_up = nullptr;

  // boundary (single-scene contexts)
  int nb = 0, nbodies = 0;
  DBuf<float4> BL, BP, BPS, FP, FS, FB;
  DBuf<int> BODY, BODY_PTR, BODY_IDX;
  DBuf<float> PSI;
  DBuf<float> POSE;
  DBuf<uint8_t> STATIC;
  DBuf<float> OUT6;
  DBuf<int> Acstart, ACLS, AOFF, ATPERM, ATKEY, APERM, AMPERM, AMKEY;
  DBuf<uint8_t> ABIGB;
  DBuf<int> AHID, AHIDP, AHCNT, AHUGE;  // index-space rank of the huge classes (per class, per key,
  DBuf<uint8_t> AHMARK;                 // per class and key tile, the count and ranges)
  DBuf<int> aux_clsmap;

  // batched coupling: each scene's boundary particles occupy a segment of a
  // boundary pool, placed when the boundary is set and released with the
  // scene, so admitting or retiring a scene uploads only its own particles.
  // The body table (the bodies of the scenes with boundary particles, by
  // ascending scene id, each scene's bodies in its own order) is rebuilt
  // from the per-scene records when that set changes: O(scenes + bodies).
  struct SceneBnd {
    std::vector<float> local3, psi;
    std::vector<int> body;
    std::vector<uint8_t> is_static;
    int nbodies = 0;
    int off = -1;                   // pool segment [off, off + body.size())
    std::vector<int> bstart, bend;  // per body: its particles' range in the segment's index list
  };
  std::vector<SceneBnd> sbnd;
  RangeAlloc bnd_alloc;
  int bnd_cap = 0;
  bool body_dirty = false;
  int bb_n = 0;       // pool positions in use or released (the kernels' range)
  int bb_live = 0;    // placed boundary particles
  int bb_bodies = 0;  // bodies in the body table
  struct PointSort {
    DBuf<int> cst, cls, off, tperm, tkey, perm, mperm, mkey;
    DBuf<uint8_t> bigb;
    DBuf<int> hid, hidp, hcnt, huge;  // index-space rank of the huge classes
    DBuf<uint8_t> hmark;
  };

  // Index-space rank of an auxiliary sort's huge classes: the buffers for
  // ncls classes and keys unique in [0, nkeys), cleared before the scatter.
  struct HugeBufs {
    int* hid;  // indexed by class id (offset like the class starts)
    int *hidp, *hcnt, *huge;
    uint8_t* hmark;
    int nkeys;
  };
  int huge_prepare(DBuf<int>& hid, DBuf<int>& hidp, DBuf<int>& hcnt, DBuf<int>& huge, DBuf<uint8_t>& hmark, int ncls,
                   int nkeys, int cls_base, HugeBufs& hb) {
    const int words = cdiv(std::max(nkeys, 1), 4);
    NPX_HIP(hid.ensure((size_t)std::max(ncls, 1)));
    NPX_HIP(hidp.ensure((size_t)std::max(nkeys, 1)));
    NPX_HIP(hcnt.ensure((size_t)NPX_HUGE_MAX * cdiv(std::max(nkeys, 1), NPX_BLOCK)));
    NPX_HIP(huge.ensure(1 + 3 * NPX_HUGE_MAX));
    NPX_HIP(hmark.ensure((size_t)words * 4));
    NPX_TRY(launch("huge_reset", k_huge_reset, cdiv(words, 256), 256, (unsigned*)hmark.p, words, huge.p));
    hb = HugeBufs{hid.p - cls_base, hidp.p, hcnt.p, huge.p, hmark.p, nkeys};
    return NPX_OK;
  }

  // Rank inside classes of an auxiliary sort (k_aux_rank_sort, k_aux_rank,
  // the index-space rank of huge classes and the merge launches of the other
  // large classes): off is reused for each sorted slot's class, mperm / mkey
  // are the merge launches' other buffers, bigb one byte per 256-slot block.
  // hb == nullptr: no index-space rank (every large class merges).
  int aux_rank_all(const char* tag, int n, int ncls, const int* cst, const int* cls, int* off, int* tperm, int* tkey,
                   int* mperm, int* mkey, uint8_t* bigb, const float4* P, int* perm, float4* PS,
                   const HugeBufs* hb) {
    const int nblk = cdiv(n, NPX_BLOCK);
    const int hl = hb ? huge_levels : 31;
    const int* hid = hb ? hb->hid : nullptr;
    NPX_TRY(launch(tag, k_aux_rank_sort, nblk, NPX_BLOCK, ncls, cst, cls, tperm, tkey, bigb, hid, hl));
    NPX_TRY(launch(tag, k_aux_rank, nblk, NPX_BLOCK, n, ncls, cst, cls, (const int*)tperm, (const int*)tkey, P, perm,
                   PS, off, hid, hl, hb ? hb->hmark : nullptr, hb ? hb->hidp : nullptr));
    if (hb) {
      const int ntiles = cdiv(std::max(hb->nkeys, 1), NPX_BLOCK);
      NPX_TRY(launch("huge_hist", k_aux_huge_hist, ntiles, NPX_BLOCK, hb->nkeys, (const int*)hb->huge,
                     (const uint8_t*)hb->hmark, ntiles, hb->hcnt));
      NPX_TRY(launch("huge_scan", k_huge_scan, NPX_HUGE_MAX, 1024, (const SceneC*)nullptr, (const int*)hb->huge,
                     ntiles, hb->hcnt));
      NPX_TRY(launch("huge_place", k_aux_huge_place, ntiles, NPX_BLOCK, hb->nkeys, (const int*)hb->huge,
                     (const uint8_t*)hb->hmark, (const int*)hb->hidp, ntiles, (const int*)hb->hcnt, P, perm, PS));
    }
    int levels = 0;
    while ((1 << levels) < nblk + 1) ++levels;
    for (int L = 0; L < levels; L += NPX_MERGE_LOG) {
      const bool even = ((L / NPX_MERGE_LOG) & 1) == 0;
      NPX_TRY(launch(tag, k_aux_rank_merge, nblk, NPX_BLOCK, ncls, cst, (const int*)off, (const uint8_t*)bigb, L,
                     (const int*)(even ? tperm : mperm), (const int*)(even ? tkey : mkey), even ? mperm : tperm,
                     even ? mkey : tkey, P, perm, PS, hid, hl));
    }
    return NPX_OK;
  }
  PointSort bsort, fsort;
  DBuf<float4> BBL, BBP, BBPS, BFP, BFS, BFB;
  DBuf<int> BBODY, BSCN, BBIDX, FPSC;     // pool: local body, scene (-1: free), per-body index lists
  DBuf<int> BFK;  // per slot: the fluid sort's key (segment offset + original index)
  DBuf<int> BBSTART, BBEND, BSBASE;       // body table: index-list ranges; per scene its first body
  DBuf<float> BPSI, BPOSE, BOUT6;
  DBuf<uint8_t> BSTAT;

  // Sorts points of known scene (pscene, -1: none) into the global classes:
  // ps.cst holds the class starts, PS the payload in class order, ascending
  // key inside a class (key = point index when NULL).
  // key: NULL (the point index) or keys unique in [0, n) over all points.
  int sort_scene_points(PointSort& ps, const float4* P, const int* pscene, int n, const int* key, float4* PS) {
    const int ncls = total_cls();
    NPX_HIP(ps.cst.ensure((size_t)ncls + 1));
    const size_t m = (size_t)std::max(n, 1);
    NPX_HIP(ps.cls.ensure(m));
    NPX_HIP(ps.off.ensure(m));
    NPX_HIP(ps.tperm.ensure(m));
    NPX_HIP(ps.tkey.ensure(m));
    NPX_HIP(ps.perm.ensure(m));
    NPX_HIP(ps.mperm.ensure(m));
    NPX_HIP(ps.mkey.ensure(m));
    NPX_HIP(ps.bigb.ensure(m / NPX_BLOCK + 2));
    NPX_TRY(launch("bsort_fill", k_fill_i32, cdiv(ncls + 1, 256), 256, ps.cst.p, 0, ncls + 1));
    if (n == 0) return NPX_OK;
    NPX_TRY(launch("bsort_bin", k_bin_scene_points, cdiv(n, NPX_BIN_BLOCK), NPX_BIN_BLOCK, (const SceneC*)d_scenes.p,
                   P, pscene, n,
                   (const int*)clsmap.p, ps.cst.p, ps.cls.p, ps.off.p));
    NPX_TRY(scan(ps.cst.p, ncls + 1));
    HugeBufs hb;
    NPX_TRY(huge_prepare(ps.hid, ps.hidp, ps.hcnt, ps.huge, ps.hmark, ncls, n, 0, hb));
    NPX_TRY(launch("bsort_scatter", k_aux_scatter, cdiv(n, 256), 256, n, (const int*)ps.cst.p, (const int*)ps.cls.p,
                   (const int*)ps.off.p, key, ps.tperm.p, ps.tkey.p, hb.hid, hb.huge, huge_levels));
    return aux_rank_all("bsort_rank", n, ncls, ps.cst.p, ps.cls.p, ps.off.p, ps.tperm.p, ps.tkey.p, ps.mperm.p,
                        ps.mkey.p, ps.bigb.p, P, ps.perm.p, PS, &hb);
  }

  // Runs of identical points of a sorted payload (k_payload_runs /
  // k_payload_pieces) into CRPI / CRPS: n sorted slots at most, cst the class
  // starts (ncls + 1 entries, by class id), tcls each sorted slot's class.
  int payload_runs(int n, const int* cst, int ncls, const int* tcls, const float4* PS) {
    NPX_HIP(CRF.ensure((size_t)n + 1));
    NPX_HIP(CRPI.ensure((size_t)std::max(n, 1)));
    NPX_HIP(CRPS.ensure((size_t)std::max(n, 1)));
    NPX_TRY(launch("payload_runs", k_payload_runs, cdiv(n + 1, 256), 256, n, cst, ncls, tcls, PS, CRF.p));
    NPX_TRY(scan(CRF.p, n + 1));
    NPX_TRY(launch("payload_pieces", k_payload_pieces, cdiv(n, 256), 256, n, cst, ncls, (const int*)CRF.p, CRPI.p,
                   CRPS.p));
    return NPX_OK;
  }

  // Uploads scene s's boundary into its pool segment: local positions, local
  // body (-1: world-owned), scene, psi, and the segment's index list (pool
  // positions of the particles of each body, ascending, body by body).
  int upload_boundary_segment(int s) {
    SceneBnd& B = sbnd[s];
    const int nb = (int)B.body.size(), off = B.off;
    std::vector<float4> bl(nb);
    std::vector<int> lb(nb), sc(nb, s), idx;
    idx.reserve(nb);
    for (int b = 0; b < nb; ++b) {
      bl[b] = make_float4(B.local3[3 * b], B.local3[3 * b + 1], B.local3[3 * b + 2], 0.0f);
      const int k = B.body[b];
      lb[b] = (k >= 0 && k < B.nbodies) ? k : -1;
    }
    B.bstart.assign(B.nbodies, 0);
    B.bend.assign(B.nbodies, 0);
    std::vector<int> cnt(B.nbodies + 1, 0);
    for (int b = 0; b < nb; ++b)
      if (lb[b] >= 0) cnt[lb[b] + 1]++;
    for (int k = 0; k < B.nbodies; ++k) cnt[k + 1] += cnt[k];
    idx.assign(std::max(cnt[B.nbodies], 1), 0);
    std::vector<int> fillp(cnt.begin(), cnt.end());
    for (int b = 0; b < nb; ++b)
      if (lb[b] >= 0) idx[fillp[lb[b]]++] = off + b;
    for (int k = 0; k < B.nbodies; ++k) {
      B.bstart[k] = cnt[k];
      B.bend[k] = cnt[k + 1];
    }
    NPX_TRY(copy_sync(BBL.p + off, bl.data(), sizeof(float4) * nb, hipMemcpyHostToDevice));
    NPX_TRY(copy_sync(BBODY.p + off, lb.data(), sizeof(int) * nb, hipMemcpyHostToDevice));
    NPX_TRY(copy_sync(BSCN.p + off, sc.data(), sizeof(int) * nb, hipMemcpyHostToDevice));
    NPX_TRY(copy_sync(BPSI.p + off, B.psi.data(), sizeof(float) * nb, hipMemcpyHostToDevice));
    if (cnt[B.nbodies] > 0)
      NPX_TRY(copy_sync(BBIDX.p + off, idx.data(), sizeof(int) * cnt[B.nbodies], hipMemcpyHostToDevice));
    bb_n = std::max(bb_n, off + nb);
    return NPX_OK;
  }

  // Pool arrays of capacity cap, every position free (scene -1).
  int alloc_boundary_pool(int cap) {
    const size_t n = (size_t)std::max(cap, 1);
    for (auto* b : {&BBL, &BBP, &BBPS, &BFB}) {
      b->release();
      NPX_HIP(b->ensure(n));
    }
    for (auto* b : {&BBODY, &BSCN, &BBIDX}) {
      b->release();
      NPX_HIP(b->ensure(n));
    }
    BPSI.release();
    NPX_HIP(BPSI.ensure(n));
    NPX_TRY(launch("bnd_free", k_fill_i32, cdiv((int)n, 256), 256, BSCN.p, -1, (int)n));
    bnd_cap = cap;
    bnd_alloc.init(cap);
    bb_n = 0;
    return NPX_OK;
  }

  // Places scene s's boundary in a pool segment and uploads it; when no
  // segment fits, the pool grows (at least doubling) and every placed scene
  // is uploaded again.
  int place_boundary(int s) {
    SceneBnd& B = sbnd[s];
    const int nb = (int)B.body.size();
    B.off = -1;
    if (nb == 0) return NPX_OK;
    int off = bnd_alloc.alloc(nb);
    if (off < 0) {
      const long long want = std::max<long long>({2ll * bnd_cap, (long long)bb_live + nb, 1 << 16});
      if (want > (1ll << 30)) return fail(NPX_ENOMEM, "boundary pool beyond 2^30 particles");
      NPX_TRY(alloc_boundary_pool((int)want));
      for (int t = 0; t < (int)sbnd.size(); ++t) {
        if (t == s || sbnd[t].off < 0) continue;
        sbnd[t].off = bnd_alloc.alloc((int)sbnd[t].body.size());
        NPX_TRY(upload_boundary_segment(t));
      }
      off = bnd_alloc.alloc(nb);
      if (off < 0) return fail(NPX_ENOMEM, "boundary pool allocation failed");
    }
    B.off = off;
    bb_live += nb;
    body_dirty = true;
    return upload_boundary_segment(s);
  }

  // Frees scene s's pool segment (its positions read as scene -1).
  int release_boundary(int s) {
    if (s >= (int)sbnd.size()) return NPX_OK;
    SceneBnd& B = sbnd[s];
    const int nb = (int)B.body.size();
    if (B.off >= 0 && nb > 0) {
      NPX_TRY(launch("bnd_free", k_fill_i32, cdiv(nb, 256), 256, BSCN.p + B.off, -1, nb));
      bnd_alloc.release(B.off, nb);
      bb_live -= nb;
      body_dirty = true;
    }
    B = SceneBnd();
    return NPX_OK;
  }

  // The body table: per scene its first body (BSBASE), per body its index
  // list range (BBSTART / BBEND) and static flag.
  int rebuild_bodies() {
    std::vector<int> base(max_scenes, 0), bs, be;
    std::vector<uint8_t> st;
    int total = 0;
    for (int s = 0; s < max_scenes && s < (int)sbnd.size(); ++s) {
      const SceneBnd& B = sbnd[s];
      if (!scenes[s].used || B.off < 0) continue;
      base[s] = total;
      for (int k = 0; k < B.nbodies; ++k) {
        bs.push_back(B.off + B.bstart[k]);
        be.push_back(B.off + B.bend[k]);
        st.push_back(B.is_static[k]);
      }
      total += B.nbodies;
    }
    bb_bodies = total;
    const size_t nbod = (size_t)std::max(total, 1);
    NPX_HIP(BSBASE.ensure((size_t)max_scenes));
    NPX_HIP(BBSTART.ensure(nbod));
    NPX_HIP(BBEND.ensure(nbod));
    NPX_HIP(BPOSE.ensure(nbod * 7));
    NPX_HIP(BOUT6.ensure(nbod * 6));
    NPX_HIP(BSTAT.ensure(nbod));
    NPX_TRY(copy_sync(BSBASE.p, base.data(), sizeof(int) * max_scenes, hipMemcpyHostToDevice));
    if (total > 0) {
      NPX_TRY(copy_sync(BBSTART.p, bs.data(), sizeof(int) * total, hipMemcpyHostToDevice));
      NPX_TRY(copy_sync(BBEND.p, be.data(), sizeof(int) * total, hipMemcpyHostToDevice));
      NPX_TRY(copy_sync(BSTAT.p, st.data(), total, hipMemcpyHostToDevice));
    }
    body_dirty = false;
    return NPX_OK;
  }

  // Boundary positions, boundary density and coupling forces of every scene
  // with a boundary (FluidWorld::step after the solve, fluid_world.cpp:58-82),
  // out6: per body force and torque.
  int batch_coupling(const float* pose7, int nbodies, float* out6) {
    if (body_dirty) NPX_TRY(rebuild_bodies());
    if (nbodies != bb_bodies) return fail(NPX_EINVAL, "expected %d body poses, got %d", bb_bodies, nbodies);
    if (bb_live == 0) return NPX_OK;
    if (nbodies > 0 && (!pose7 || !out6)) return fail(NPX_EINVAL, "pose7 / out6 is NULL");
    for (const auto& H : scenes)
      if (H.used && H.nbnd > 0 && !H.window_valid) return fail(NPX_ESTATE, "no step has run yet");
    NPX_TRY(upload_bodies(BPOSE.p, pose7, nullptr, nullptr, nbodies));
    NPX_TRY(launch("bpos", k_bpos_b, cdiv(bb_n, 256), 256, (const float4*)BBL.p, (const int*)BBODY.p,
                   (const int*)BSCN.p, (const int*)BSBASE.p, (const float*)BPSI.p, bb_n, (const float*)BPOSE.p,
                   nbodies, BBP.p));
    NPX_TRY(sort_scene_points(bsort, BBP.p, BSCN.p, bb_n, nullptr, BBPS.p));
    const SceneC* sc = d_scenes.p;
    const int* bs = d_blk_scene.p;
    NPX_TRY(launch("boundary_density", k_boundary_density_b, hw_blocks, NPX_BLOCK, sc, bs, (const float4*)X.p,
                   (const int*)clsmap.p, (const int*)bsort.cst.p, (const float4*)BBPS.p, RHO.p, d_flags.p));
    NPX_HIP(BFP.ensure(slot_cap));
    NPX_HIP(BFS.ensure(slot_cap));
    NPX_HIP(FPSC.ensure(slot_cap));
    NPX_HIP(BFK.ensure(slot_cap));
    NPX_TRY(launch("fluid_payload", k_fluid_payload, hw_blocks, NPX_BLOCK, sc, bs, (const float4*)X.p,
                   (const float*)RHO.p, BFP.p, (const int*)ID.p, BFK.p));
    NPX_TRY(launch("coupled_slots", k_coupled_slots, hw_blocks, NPX_BLOCK, sc, bs, FPSC.p));
    // keys: segment offset + index, unique over the batch
    NPX_TRY(sort_scene_points(fsort, BFP.p, FPSC.p, hw_blocks * NPX_BLOCK, BFK.p, BFS.p));
    NPX_TRY(payload_runs(hw_blocks * NPX_BLOCK, fsort.cst.p, total_cls(), fsort.off.p, BFS.p));
    // a wave per boundary particle
    NPX_TRY(launch("coupling", k_coupling_b, cdiv(bb_n, 4), 256, sc, (const float4*)BBP.p, (const int*)BSCN.p,
                   (const int*)BBODY.p, (const int*)BSBASE.p, bb_n, (const uint8_t*)BSTAT.p, nbodies,
                   (const int*)clsmap.p, (const int*)fsort.cst.p, (const float4*)BFS.p, (const int*)CRPI.p,
                   (const int*)CRPS.p, BFB.p));
    if (nbodies > 0) {
      NPX_TRY(launch("body_reduce", k_body_reduce, nbodies, 64, (const int*)BBSTART.p, (const int*)BBEND.p,
                     (const int*)BBIDX.p, nbodies, (const float4*)BFB.p, (const float4*)BBP.p, (const float*)BPOSE.p,
                     BOUT6.p));
      NPX_TRY(read_bodies(out6, BOUT6.p, nbodies));
    }
    return NPX_OK;  // a scene whose particles left its window is reported by npx_scene_status
  }

  // timing
  bool timing = false;
  std::vector<KTimer> timers;
  std::vector<hipEvent_t> ev_pool;

  int fail(int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    err = buf;
    return code;
  }

  hipEvent_t get_event() {
    if (!ev_pool.empty()) {
      hipEvent_t e = ev_pool.back();
      ev_pool.pop_back();
      return e;
    }
    hipEvent_t e;
    (void)hipEventCreate(&e);
    return e;
  }

  template <class K, class... A>
  int launch(const char* name, K kern, int blocks, int threads, A... args) {
    return launch_shm(name, kern, blocks, threads, 0, args...);
  }

  // Launch-shape limits: the device's LDS per workgroup (queried at init) and,
  // per kernel, its static LDS and largest block size (its launch bounds),
  // looked up once and cached.
  size_t max_lds = 65536;
  struct KernLimits {
    size_t static_lds;
    int max_threads;
  };
  std::map<const void*, KernLimits> klimits;

  // Refuses a launch the device could not run as asked: a negative grid, a
  // block size that is not a multiple of 64 in [64, 1024] or above the
  // kernel's launch bounds, or static plus dynamic LDS above the device's
  // per-workgroup limit.  Such a launch is an error of this library; it is
  // reported instead of being handed to the queue.
  int check_launch(const char* name, const void* kern, int blocks, int threads, size_t shmem) {
    if (blocks < 0) return fail(NPX_EINVAL, "launch %s: negative grid (%d blocks)", name, blocks);
    if (threads < 64 || threads > 1024 || (threads & 63) != 0)
      return fail(NPX_EINVAL, "launch %s: block size %d is not a multiple of 64 in [64, 1024]", name, threads);
    auto it = klimits.find(kern);
    if (it == klimits.end()) {
      hipFuncAttributes a;
      NPX_HIP(hipFuncGetAttributes(&a, kern));
      it = klimits.emplace(kern, KernLimits{(size_t)a.sharedSizeBytes, a.maxThreadsPerBlock}).first;
    }
    if (threads > it->second.max_threads)
      return fail(NPX_EINVAL, "launch %s: block size %d exceeds the kernel's launch bounds (%d)", name, threads,
                  it->second.max_threads);
    if (it->second.static_lds + shmem > max_lds)
      return fail(NPX_EINVAL, "launch %s: %zu B static + %zu B dynamic LDS exceed the %zu B per workgroup", name,
                  it->second.static_lds, shmem, max_lds);
    return NPX_OK;
  }

  template <class K, class... A>
  int launch_shm(const char* name, K kern, int blocks, int threads, size_t shmem, A... args) {
    NPX_TRY(check_launch(name, (const void*)kern, blocks, threads, shmem));
    if (blocks == 0) return NPX_OK;
    KTimer t;
    if (timing) {
      t.name = name;
      t.a = get_event();
      t.b = get_event();
      NPX_HIP(hipEventRecord(t.a, stream));
    }
    hipLaunchKernelGGL(kern, dim3(blocks), dim3(threads), shmem, stream, args...);
    NPX_HIP(hipGetLastError());
    if (timing) {
      NPX_HIP(hipEventRecord(t.b, stream));
      timers.push_back(t);
    }
    return NPX_OK;
  }

  // Stream-ordered blocking copy: every kernel runs on `stream` (non-blocking
  // with respect to the null stream), so copies must be ordered on it too.
  int copy_sync(void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
    if (bytes == 0) return NPX_OK;
    NPX_HIP(hipMemcpyAsync(dst, src, bytes, kind, stream));
    NPX_HIP(hipStreamSynchronize(stream));
    return NPX_OK;
  }

  // A scene that (re)starts has no list statistics: its first step builds lists.
  int reset_list_stats(int sid) {
    static const int zero[NPX_LST_STRIDE] = {};
    return copy_sync(LST.p + NPX_LST_STRIDE * (size_t)sid, zero, sizeof(zero), hipMemcpyHostToDevice);
  }

  int stage_bodies(int nbodies) {
    if (nbodies <= h_body_cap) return NPX_OK;
    NPX_HIP(hipStreamSynchronize(stream));
    if (h_pose) NPX_HIP(hipHostFree(h_pose));
    if (h_static) NPX_HIP(hipHostFree(h_static));
    if (h_out) NPX_HIP(hipHostFree(h_out));
    h_pose = nullptr;
    h_static = nullptr;
    h_out = nullptr;
    h_body_cap = 0;
    const int cap = std::max(nbodies, 64);
    NPX_HIP(hipHostMalloc((void**)&h_pose, sizeof(float) * 7 * cap, hipHostMallocDefault));
    NPX_HIP(hipHostMalloc((void**)&h_static, (size_t)cap, hipHostMallocDefault));
    NPX_HIP(hipHostMalloc((void**)&h_out, sizeof(float) * 6 * cap, hipHostMallocDefault));
    if (!ev_up) NPX_HIP(hipEventCreateWithFlags(&ev_up, hipEventDisableTiming));
    h_body_cap = cap;
    return NPX_OK;
  }

  // Asynchronous upload of n bodies' poses (and is_static flags when given)
  // through the pinned staging buffers.
  int upload_bodies(float* d_pose, const float* pose7, uint8_t* d_static, const uint8_t* is_static, int n) {
    if (n <= 0) return NPX_OK;
    NPX_TRY(stage_bodies(n));
    NPX_HIP(hipEventSynchronize(ev_up));  // the previous upload has left the staging buffers
    std::memcpy(h_pose, pose7, sizeof(float) * 7 * n);
    NPX_HIP(hipMemcpyAsync(d_pose, h_pose, sizeof(float) * 7 * n, hipMemcpyHostToDevice, stream));
    if (is_static) {
      std::memcpy(h_static, is_static, (size_t)n);
      NPX_HIP(hipMemcpyAsync(d_static, h_static, (size_t)n, hipMemcpyHostToDevice, stream));
    }
    NPX_HIP(hipEventRecord(ev_up, stream));
    return NPX_OK;
  }

  // Blocking readback of n bodies' force + torque (6 floats each) through the
  // pinned staging buffer.
  int read_bodies(float* out6, const float* d_out6, int n) {
    if (n <= 0) return NPX_OK;
    NPX_TRY(stage_bodies(n));
    NPX_HIP(hipMemcpyAsync(h_out, d_out6, sizeof(float) * 6 * n, hipMemcpyDeviceToHost, stream));
    NPX_HIP(hipStreamSynchronize(stream));
    std::memcpy(out6, h_out, sizeof(float) * 6 * n);
    return NPX_OK;
  }

  // ---------------------------------------------------------------- setup
  int init(int slots, int cells, int nscenes, uint32_t fl, bool is_batch) {
    flags = fl;
    fast = (fl & NPX_FLAG_FAST_MATH) != 0;
    batch = is_batch;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(NPX_EDEVICE, "no HIP device available");
    NPX_HIP(hipGetDevice(&device));
    int lds = 0;
    NPX_HIP(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
    if (lds > 0) max_lds = (size_t)lds;
    NPX_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    max_scenes = std::max(1, nscenes);
    scenes.assign(max_scenes, SceneH());
    NPX_HIP(d_scenes.ensure(max_scenes));
    NPX_HIP(d_flags.ensure(1 + max_scenes));
    NPX_HIP(LST.ensure(NPX_LST_STRIDE * (size_t)max_scenes));
    // no list statistics yet (under list policy 1 the first step builds lists)
    NPX_HIP(hipMemsetAsync(LST.p, 0, sizeof(int) * NPX_LST_STRIDE * (size_t)max_scenes, stream));
    NPX_HIP(hipHostMalloc((void**)&h_flags, sizeof(int) * (1 + max_scenes), hipHostMallocDefault));
    NPX_HIP(hipMemsetAsync(d_flags.p, 0, sizeof(int) * (1 + max_scenes), stream));
    NPX_TRY(reserve_slots(slots));
    cls_cap = 0;
    if (cells > 0) NPX_TRY(reserve_cells(cells));
    slot_alloc.init(slot_cap / NPX_BLOCK);
    cls_alloc.init(cls_cap);
    return NPX_OK;
  }

  int reserve_slots(int slots) {
    if (slots > NPX_MAX_SLOTS) return fail(NPX_EINVAL, "%d particle slots requested; at most %d", slots, NPX_MAX_SLOTS);
    int cap = cdiv(std::max(slots, 1), NPX_BLOCK) * NPX_BLOCK;
    if (cap <= slot_cap) return NPX_OK;
    if (slot_cap > 0 && batch) return fail(NPX_ENOMEM, "batch slot capacity is fixed at creation");
    size_t n = (size_t)cap;
    NPX_HIP(X.ensure(n));
    NPX_HIP(X2.ensure(n));
    NPX_HIP(V.ensure(n));
    NPX_HIP(V2.ensure(n));
    NPX_HIP(XS0.ensure(n + NPX_SLOT_PAD));  // test_slots reads up to three slots past a range
    NPX_HIP(XS1.ensure(n + NPX_SLOT_PAD));
    NPX_HIP(VT.ensure(n));
    NPX_HIP(OM.ensure(n));
    NPX_HIP(VX.ensure(n));
    NPX_HIP(RHO.ensure(n));
    NPX_HIP(RHO2.ensure(n));
    NPX_HIP(LAM.ensure(n));
    NPX_HIP(LAM2.ensure(n));
    NPX_HIP(ID.ensure(n));
    NPX_HIP(ID2.ensure(n));
    NPX_HIP(CELL.ensure(n));
    NPX_HIP(CELL2.ensure(n));
    NPX_HIP(CLS.ensure(n));
    NPX_HIP(OFF.ensure(n));
    NPX_HIP(TPERM.ensure(n));
    NPX_HIP(TKEY.ensure(n));
    NPX_HIP(PERM.ensure(n));
    NPX_HIP(MPERM.ensure(n));
    NPX_HIP(MKEY.ensure(n));
    NPX_HIP(BIGB.ensure(n / NPX_BLOCK + 4));
    NPX_HIP(HMARK.ensure(n));
    NPX_HIP(HIDP.ensure(n));
    NPX_HIP(HCNT.ensure((size_t)NPX_HUGE_MAX * (n / NPX_BLOCK)));
    NPX_HIP(HUGE.ensure(1 + 3 * NPX_HUGE_MAX));
    NPX_HIP(XORD.ensure(1));
    NPX_HIP(hipMemsetAsync(XORD.p, 0, sizeof(int), stream));  // k_pieces sets it every step
    NPX_HIP(SKL.ensure(n * NPX_SKIN_CAP));
    NPX_HIP(SKN.ensure(n));
    NPX_HIP(EXL.ensure(n * NPX_SKIN_CAP));
    NPX_HIP(EXN.ensure(n));
    NPX_HIP(CDB.ensure(n));
    NPX_HIP(RSTART.ensure(n));
    NPX_HIP(PI.ensure(n));
    NPX_HIP(PS.ensure(n));
    NPX_HIP(BOFF.ensure(n / NPX_BLOCK + 5));
    NPX_HIP(DL.ensure(n));
    NPX_HIP(DN.ensure(1));
    // +4 entries of -1: neighbour passes with blocks larger than a segment
    // may run past the last segment
    NPX_HIP(d_blk_scene.ensure(n / NPX_BLOCK + 4));
    slot_cap = cap;
    h_blk_scene.assign(slot_cap / NPX_BLOCK + 4, -1);
    return NPX_OK;
  }

  int reserve_cells(int cells) {
    if (cells <= cls_cap) return NPX_OK;
    NPX_HIP(clsmap.ensure((size_t)cells));
    NPX_HIP(cstart.ensure((size_t)cells + 1));
    NPX_HIP(CMAX.ensure(2 * (size_t)cells + (size_t)max_scenes));
    NPX_HIP(bsum.ensure((size_t)cdiv(cells + 1, 1024) + 1));
    cls_cap = cells;
    return NPX_OK;
  }

  int scan(int* data, int n) {
    const int nb = cdiv(n, 1024);
    NPX_HIP(bsum.ensure((size_t)nb + 1));
    NPX_TRY(launch("scan_reduce", k_scan_reduce, nb, 256, (const int*)data, n, bsum.p));
    NPX_TRY(launch("scan_bsum", k_scan_bsum, 1, 1024, bsum.p, nb));
    NPX_TRY(launch("scan_down", k_scan_down, nb, 256, data, n, (const int*)bsum.p));
    return NPX_OK;
  }

  // Builds the class map of a window into out[0..ncell) with values cls_base + class.
  int build_class_table(const int wlo[3], const int wdim[3], int cls_base, int* out) {
    const long long ncell = (long long)wdim[0] * wdim[1] * wdim[2];
    unsigned cap = 1;
    while ((long long)cap < 2 * ncell) cap <<= 1;
    NPX_HIP(htk.ensure(cap));
    NPX_HIP(htv.ensure(cap));
    const int nc = (int)ncell;
    NPX_TRY(launch("ht_clear", k_ht_clear, cdiv(cap, 256), 256, htk.p, htv.p, (int)cap));
    NPX_TRY(launch("ht_insert", k_ht_insert, cdiv(nc, 256), 256, htk.p, htv.p, cap - 1, wlo[0], wlo[1], wlo[2],
                   wdim[0], wdim[1], nc));
    NPX_TRY(launch("ht_lookup", k_ht_lookup, cdiv(nc, 256), 256, (const unsigned long long*)htk.p,
                   (const int*)htv.p, cap - 1, wlo[0], wlo[1], wlo[2], wdim[0], wdim[1], nc, cls_base, out));
    return NPX_OK;
  }

  static void window_from_bounds(float inv_h, const float lo[3], const float hi[3], int margin, int wlo[3],
                                 int wdim[3]) {
    for (int d = 0; d < 3; ++d) {
      int a = (int)std::floor(lo[d] * inv_h) - margin;
      int b = (int)std::floor(hi[d] * inv_h) + margin;
      wlo[d] = a;
      wdim[d] = b - a + 1;
    }
  }

  int upload_tables() {
    if (!table_dirty) return NPX_OK;
    std::vector<SceneC> sc(max_scenes);
    for (int s = 0; s < max_scenes; ++s) {
      if (scenes[s].used)
        sc[s] = make_scene_consts(scenes[s], skin_frac);
      else
        std::memset(&sc[s], 0, sizeof(SceneC));
    }
    NPX_HIP(hipMemcpyAsync(d_scenes.p, sc.data(), sizeof(SceneC) * max_scenes, hipMemcpyHostToDevice, stream));
    NPX_HIP(hipMemcpyAsync(d_blk_scene.p, h_blk_scene.data(), sizeof(int) * h_blk_scene.size(),
                           hipMemcpyHostToDevice, stream));
    // the host vectors must stay alive until the copies are done
    NPX_HIP(hipStreamSynchronize(stream));
    hw_blocks = 0;
    for (int b = 0; b < slot_cap / NPX_BLOCK; ++b)
      if (h_blk_scene[b] >= 0) hw_blocks = b + 1;
    hw_cls = 0;
    for (const auto& H : scenes)
      if (H.used && H.window_valid) hw_cls = std::max(hw_cls, H.cls_base + H.ncell);
    table_dirty = false;
    return NPX_OK;
  }

  // --------------------------------------------------------- scenes
  int scene_window(SceneH& H) {
    if (!H.prm.use_domain_bounds) return NPX_OK;  // set per step from the particle bounds
    int wlo[3], wdim[3];
    window_from_bounds(1.0f / H.prm.kernel_radius, H.prm.domain_lower, H.prm.domain_upper, NPX_WINDOW_MARGIN,
                       wlo, wdim);
    return set_window(H, wlo, wdim);
  }

  int set_window(SceneH& H, const int wlo[3], const int wdim[3]) {
    const long long nc = (long long)wdim[0] * wdim[1] * wdim[2];
    if (wdim[0] < 3 || wdim[1] < 3 || wdim[2] < 3 || nc > (1ll << 30))
      return fail(NPX_EDOMAIN, "cell window %dx%dx%d is invalid or too large", wdim[0], wdim[1], wdim[2]);
    if (H.window_valid && H.ncell == nc && std::equal(wlo, wlo + 3, H.wlo) && std::equal(wdim, wdim + 3, H.wdim))
      return NPX_OK;
    if (!batch) {
      // single scene: the class range is the whole table
      NPX_TRY(reserve_cells((int)nc));
      H.cls_base = 0;
    } else {
      if (H.window_valid) cls_alloc.release(H.cls_base, H.ncell);
      int b = cls_alloc.alloc((int)nc);
      if (b < 0) return fail(NPX_ENOMEM, "cell capacity exhausted (%lld cells requested)", nc);
      H.cls_base = b;
    }
    H.ncell = (int)nc;
    std::copy(wlo, wlo + 3, H.wlo);
    std::copy(wdim, wdim + 3, H.wdim);
    H.window_valid = true;
    hw_cls = std::max(hw_cls, H.cls_base + H.ncell);
    table_dirty = true;
    NPX_TRY(build_class_table(H.wlo, H.wdim, H.cls_base, clsmap.p + H.cls_base));
    return NPX_OK;
  }

  int upload_particles(SceneH& H, const float* pos3, const float* vel3) {
    const int n = H.n;
    std::vector<float4> x(n), v(n);
    std::vector<int> id(n);
    for (int i = 0; i < n; ++i) {
      x[i] = make_float4(pos3[3 * i], pos3[3 * i + 1], pos3[3 * i + 2], 0.0f);
      v[i] = vel3 ? make_float4(vel3[3 * i], vel3[3 * i + 1], vel3[3 * i + 2], 0.0f) : make_float4(0, 0, 0, 0);
      id[i] = i;
    }
    const size_t o = H.seg_off;
    if (n > 0) {
      NPX_HIP(hipMemcpyAsync(X.p + o, x.data(), sizeof(float4) * n, hipMemcpyHostToDevice, stream));
      NPX_HIP(hipMemcpyAsync(V.p + o, v.data(), sizeof(float4) * n, hipMemcpyHostToDevice, stream));
      NPX_HIP(hipMemcpyAsync(ID.p + o, id.data(), sizeof(int) * n, hipMemcpyHostToDevice, stream));
      NPX_HIP(hipMemsetAsync(RHO.p + o, 0, sizeof(float) * n, stream));
      NPX_HIP(hipMemsetAsync(LAM.p + o, 0, sizeof(float) * n, stream));
    }
    NPX_HIP(hipStreamSynchronize(stream));
    return NPX_OK;
  }

  int download(const SceneH& H, float* pos3, float* vel3, float* rho, float* lam) {
    NPX_TRY(check_flags());
    const int n = H.n;
    if (n == 0) return NPX_OK;
    std::vector<float4> x(n), v(n);
    std::vector<float> r(n), l(n);
    std::vector<int> id(n);
    const size_t o = H.seg_off;
    NPX_HIP(hipMemcpyAsync(x.data(), X.p + o, sizeof(float4) * n, hipMemcpyDeviceToHost, stream));
    NPX_HIP(hipMemcpyAsync(v.data(), V.p + o, sizeof(float4) * n, hipMemcpyDeviceToHost, stream));
    NPX_HIP(hipMemcpyAsync(r.data(), RHO.p + o, sizeof(float) * n, hipMemcpyDeviceToHost, stream));
    NPX_HIP(hipMemcpyAsync(l.data(), LAM.p + o, sizeof(float) * n, hipMemcpyDeviceToHost, stream));
    NPX_HIP(hipMemcpyAsync(id.data(), ID.p + o, sizeof(int) * n, hipMemcpyDeviceToHost, stream));
    NPX_HIP(hipStreamSynchronize(stream));
    for (int q = 0; q < n; ++q) {
      const int i = id[q];
      if (i < 0 || i >= n) return fail(NPX_ESTATE, "corrupt particle id table");
      if (pos3) {
        pos3[3 * i] = x[q].x;
        pos3[3 * i + 1] = x[q].y;
        pos3[3 * i + 2] = x[q].z;
      }
      if (vel3) {
        vel3[3 * i] = v[q].x;
        vel3[3 * i + 1] = v[q].y;
        vel3[3 * i + 2] = v[q].z;
      }
      if (rho) rho[i] = r[q];
      if (lam) lam[i] = l[q];
    }
    return NPX_OK;
  }

  int check_flags() {
    NPX_HIP(hipMemcpyAsync(h_flags, d_flags.p, sizeof(int), hipMemcpyDeviceToHost, stream));
    NPX_HIP(hipStreamSynchronize(stream));
    if (*h_flags & NPX_FLAG_WINDOW) {
      *h_flags = 0;
      NPX_HIP(hipMemsetAsync(d_flags.p, 0, sizeof(int) * (1 + max_scenes), stream));
      return fail(NPX_EDOMAIN,
                  "a particle left the cell window of its scene (more than %d cells outside the domain); "
                  "results of the affected steps are invalid",
                  NPX_WINDOW_MARGIN - 1);
    }
    return NPX_OK;
  }

  // Per-scene status bits (NPX_FLAG_*), cleared after reading.
  int scene_status(int32_t* out, int cap) {
    NPX_TRY(upload_tables());
    NPX_TRY(launch("scene_health", k_scene_health, hw_blocks, NPX_BLOCK, (const SceneC*)d_scenes.p,
                   (const int*)d_blk_scene.p, (const float4*)X.p, (const float4*)V.p, d_flags.p));
    NPX_HIP(hipMemcpyAsync(h_flags, d_flags.p, sizeof(int) * (1 + max_scenes), hipMemcpyDeviceToHost, stream));
    NPX_HIP(hipMemsetAsync(d_flags.p, 0, sizeof(int) * (1 + max_scenes), stream));
    NPX_HIP(hipStreamSynchronize(stream));
    for (int s = 0; s < std::min(cap, max_scenes); ++s) out[s] = h_flags[1 + s];
    return NPX_OK;
  }

  // Steps with a device timestamp after each one; step_ms[k] = duration of
  // step k measured between consecutive events on the context's stream.
  int step_timed(int nsteps, float* step_ms) {
    std::vector<hipEvent_t> ev(nsteps + 1);
    for (auto& e : ev) e = get_event();
    NPX_HIP(hipEventRecord(ev[0], stream));
    int r = NPX_OK;
    for (int k = 0; k < nsteps && r == NPX_OK; ++k) {
      r = step_n(1);
      if (r == NPX_OK) {
        hipError_t e = hipEventRecord(ev[k + 1], stream);
        if (e != hipSuccess) r = fail(NPX_EDEVICE, "hipEventRecord: %s", hipGetErrorString(e));
      }
    }
    if (r == NPX_OK) {
      NPX_HIP(hipStreamSynchronize(stream));
      for (int k = 0; k < nsteps; ++k) {
        float ms = 0.0f;
        NPX_HIP(hipEventElapsedTime(&ms, ev[k], ev[k + 1]));
        if (step_ms) step_ms[k] = ms;
      }
    }
    for (auto e : ev) ev_pool.push_back(e);
    return r;
  }

  // Unbounded scene: window from the bounding box of x* (host round trip).
  int fit_window_unbounded(SceneH& H) {
    const int n = H.n;
    std::vector<float4> x(n), v(n);
    const size_t o = H.seg_off;
    NPX_HIP(hipMemcpyAsync(x.data(), X.p + o, sizeof(float4) * n, hipMemcpyDeviceToHost, stream));
    NPX_HIP(hipMemcpyAsync(v.data(), V.p + o, sizeof(float4) * n, hipMemcpyDeviceToHost, stream));
    NPX_HIP(hipStreamSynchronize(stream));
    const float inv_h = 1.0f / H.prm.kernel_radius;
    const float gdt[3] = {H.g[0] * H.dt, H.g[1] * H.dt, H.g[2] * H.dt};
    int lo[3] = {0x7FFFFFFF, 0x7FFFFFFF, 0x7FFFFFFF}, hi[3] = {-0x7FFFFFFF, -0x7FFFFFFF, -0x7FFFFFFF};
    for (int i = 0; i < n; ++i) {
      const float xs[3] = {x[i].x + (v[i].x + gdt[0]) * H.dt, x[i].y + (v[i].y + gdt[1]) * H.dt,
                           x[i].z + (v[i].z + gdt[2]) * H.dt};
      for (int d = 0; d < 3; ++d) {
        const float c = std::floor(xs[d] * inv_h);
        if (!(std::fabs(c) < 1.0e8f))
          return fail(NPX_EDOMAIN, "predicted position is not finite or beyond the cell range");
        lo[d] = std::min(lo[d], (int)c);
        hi[d] = std::max(hi[d], (int)c);
      }
    }
    bool ok = H.window_valid;
    for (int d = 0; d < 3 && ok; ++d) ok = lo[d] - 1 >= H.wlo[d] && hi[d] + 1 <= H.wlo[d] + H.wdim[d] - 1;
    if (ok) return NPX_OK;
    int wlo[3], wdim[3];
    for (int d = 0; d < 3; ++d) {
      wlo[d] = lo[d] - NPX_WINDOW_MARGIN;
      wdim[d] = (hi[d] + NPX_WINDOW_MARGIN) - wlo[d] + 1;
    }
    return set_window(H, wlo, wdim);
  }

  // --------------------------------------------------------- the step
  // Classes the counting sorts and the per-class arrays cover: the single
  // scene's window, or in a batch the end of the highest class range a live
  // scene holds (hw_cls).  Work over classes (fills, scans, the displacement
  // maxima) then scales with the classes in use, not with the batch's cell
  // capacity.
  int total_cls() const {
    if (!batch) return scenes[0].used ? scenes[0].ncell : 0;
    return hw_cls;
  }
  int hw_cls = 0;  // set by upload_tables, raised by set_window

  // predict + grid: writes the cell-sorted state into the *2 buffers and x*
  // into XS0; the caller swaps the buffers to commit the reorder.
  int grid_pass(bool carry_rho) {
    const int ncls = total_cls();
    const int blocks = hw_blocks;
    NPX_TRY(launch("fill_counts", k_fill_i32, cdiv(ncls + 1, 256), 256, cstart.p, 0, ncls + 1));
    NPX_TRY(launch("predict_bin", k_predict_bin, cdiv(blocks, NPX_BIN_BLOCK / NPX_BLOCK), NPX_BIN_BLOCK,
                   (const SceneC*)d_scenes.p, (const int*)d_blk_scene.p, blocks, (const float4*)X.p, (const float4*)V.p, (const int*)clsmap.p, cstart.p,
                   CLS.p, OFF.p, CELL.p, d_flags.p, HUGE.p));
    NPX_TRY(scan(cstart.p, ncls + 1));
    // the huge classes' index h per class (read only for huge classes) lives
    // in CMAX's per-class displacement maxima until clear_cmax below
    int* hid = CMAX.p;
    NPX_TRY(launch("scatter", k_scatter, blocks, NPX_BLOCK, (const SceneC*)d_scenes.p, (const int*)d_blk_scene.p,
                   (const int*)cstart.p, (const int*)CLS.p, (const int*)OFF.p, (const int*)ID.p, TPERM.p, TKEY.p,
                   hid, HUGE.p, huge_levels));
    NPX_TRY(launch("rank_sort", k_rank_sort, blocks, NPX_BLOCK, (const SceneC*)d_scenes.p,
                   (const int*)d_blk_scene.p, (const int*)cstart.p, (const int*)CLS.p, TPERM.p, TKEY.p, BIGB.p,
                   (const int*)hid, huge_levels));
    // OFF (the binning offsets, read only by k_scatter) then holds each sorted slot's class
    NPX_TRY(launch("rank", k_rank, blocks, NPX_BLOCK, (const SceneC*)d_scenes.p, (const int*)d_blk_scene.p,
                   (const int*)cstart.p, (const int*)CLS.p, (const int*)TPERM.p, (const int*)TKEY.p, PERM.p, OFF.p,
                   (const int*)hid, HMARK.p, HIDP.p, huge_levels));
    NPX_TRY(launch("huge_hist", k_huge_hist, blocks, NPX_BLOCK, (const SceneC*)d_scenes.p, (const int*)d_blk_scene.p,
                   (const int*)HUGE.p, (const uint8_t*)HMARK.p, blocks, HCNT.p));
    NPX_TRY(launch("huge_scan", k_huge_scan, NPX_HUGE_MAX, 1024, (const SceneC*)d_scenes.p, (const int*)HUGE.p,
                   blocks, HCNT.p));
    NPX_TRY(launch("huge_place", k_huge_place, blocks, NPX_BLOCK, (const SceneC*)d_scenes.p,
                   (const int*)d_blk_scene.p, (const int*)HUGE.p, (const uint8_t*)HMARK.p, (const int*)HIDP.p, blocks,
                   (const int*)HCNT.p, PERM.p));
    // merge launches of the large classes spanning several blocks: enough for
    // a class spanning every block (a fixed count, so a step's launches do not
    // depend on the state and it can be replayed as a graph)
    int levels = 0;
    while ((1 << levels) < blocks + 1) ++levels;
    for (int L = 0; L < levels; L += NPX_MERGE_LOG) {
      const bool even = ((L / NPX_MERGE_LOG) & 1) == 0;
      NPX_TRY(launch("rank_merge", k_rank_merge, blocks, NPX_BLOCK, (const SceneC*)d_scenes.p,
                     (const int*)d_blk_scene.p, (const int*)cstart.p, (const int*)OFF.p, (const unsigned char*)BIGB.p,
                     L, (const int*)(even ? TPERM.p : MPERM.p), (const int*)(even ? TKEY.p : MKEY.p),
                     even ? MPERM.p : TPERM.p, even ? MKEY.p : TKEY.p, PERM.p, (const int*)hid, huge_levels));
    }
    // per-class displacement maxima and run flags, per-scene run flags: zero each step
    NPX_TRY(launch("clear_cmax", k_fill_i32, cdiv(2 * ncls + max_scenes, 256), 256, CMAX.p, 0, 2 * ncls + max_scenes));
    NPX_TRY(launch("gather", k_gather, blocks, NPX_BLOCK, (const SceneC*)d_scenes.p, (const int*)d_blk_scene.p,
                   (const int*)PERM.p, (const float4*)X.p, (const float4*)V.p, (const int*)ID.p, (const int*)CELL.p,
                   (const int*)clsmap.p, X2.p, V2.p, XS0.p, ID2.p, CELL2.p, fast ? 0 : 1, RSTART.p,
                   (unsigned*)CMAX.p + ncls, CMAX.p + 2 * ncls, BOFF.p));
    NPX_TRY(scan(BOFF.p, blocks + 1));
    NPX_TRY(launch("pieces", k_pieces, blocks, NPX_BLOCK, (const SceneC*)d_scenes.p, (const int*)d_blk_scene.p,
                   (const unsigned char*)RSTART.p, (const int*)BOFF.p, PI.p, PS.p, XORD.p));
    if (carry_rho) {
      NPX_TRY(launch("gather_rho", k_gather_scalar2, blocks, NPX_BLOCK, (const SceneC*)d_scenes.p,
                     (const int*)d_blk_scene.p, (const int*)PERM.p, (const float*)RHO.p, (const float*)LAM.p, RHO2.p,
                     LAM2.p));
    }
    return NPX_OK;
  }

  int prepare_step() {
    for (int s = 0; s < max_scenes; ++s) {
      SceneH& H = scenes[s];
      if (!H.used || H.n == 0) continue;
      if (!H.prm.use_domain_bounds) NPX_TRY(fit_window_unbounded(H));
      if (!H.window_valid) return fail(NPX_ESTATE, "scene %d has no cell window", s);
    }
    NPX_TRY(upload_tables());
    return NPX_OK;
  }

  int any_particles() const {
    for (const auto& H : scenes)
      if (H.used && H.n > 0) return 1;
    return 0;
  }

  // The neighbour passes of one step (persistent instantiations when PERSIST).
  template <bool PERSIST>
  int enqueue_passes(int npass, int blocks, int nd, int dense_blocks, size_t lds, const Skin& R, const BlockPlan& rm,
                     const BlockPlan& rd, int& cur) {
    const SceneC* sc = d_scenes.p;
    const int* bs = d_blk_scene.p;
    DBuf<float4>* xs[2] = {&XS0, &XS1};
    const int* cm = clsmap.p;
    const int* cst = cstart.p;
    const int* cel = CELL.p;
    for (int it = 0; it < npass; ++it) {
      auto dens = fast ? (it == 0 ? k_density_lambda<true, true, true, PERSIST> : k_density_lambda<true, false, true, PERSIST>)
                       : (it == 0 ? k_density_lambda<false, true, true, PERSIST> : k_density_lambda<false, false, true, PERSIST>);
      if (it == 0 || fast) {
        if (fast || list_policy == 1)  // (without the list policy no scene builds skin lists)
          NPX_TRY(launch(it == 0 ? (fast ? "density_lambda_build" : "density_lambda_build_lists") : "density_lambda",
                         dens, blocks, nbr_block, sc, bs, cm, cst, cel, xs[cur]->p, RHO.p, LAM.p, it, R, rm));
        if (it == 0 && !fast)  // the scenes that build no skin lists this step
          NPX_TRY(launch("density_lambda_build", k_density_lambda<false, true, false, PERSIST>, blocks, nbr_block, sc, bs, cm,
                         cst, cel, xs[cur]->p, RHO.p, LAM.p, it, R, rm));
      } else {
        NPX_TRY(launch_shm("density_lambda", dens, blocks + nd, nbr_block, lds * npx_dense_lds_floats<5>(), sc, bs,
                           cm, cst, cel, xs[cur]->p, RHO.p, LAM.p, it, R, rd));
      }
      if (!fast && it == 0)
        NPX_TRY(launch_shm("density_lambda_build_dense", k_density_dense, dense_blocks, 64,
                           sizeof(float) * npx_dense_lds_floats<5>(), sc, bs, cm, cst, cel, xs[cur]->p, RHO.p, LAM.p,
                           it, R));
      if (fast) {
        NPX_TRY(launch("delta", k_delta<true, PERSIST>, blocks, nbr_block, sc, bs, cm, cst, cel, (const float4*)xs[cur]->p,
                       xs[1 - cur]->p, it, (const float4*)X.p, (const float*)RHO.p, VT.p, R, rm));
      } else {
        NPX_TRY(launch_shm("delta", k_delta<false, PERSIST>, blocks + nd, nbr_block, lds * npx_dense_lds_floats<3>(), sc, bs,
                           cm, cst, cel, (const float4*)xs[cur]->p, xs[1 - cur]->p, it, (const float4*)X.p,
                           (const float*)RHO.p, VT.p, R, rd));
      }
      cur = 1 - cur;
    }
    NPX_TRY(launch_shm("omega_xsph", k_omega_xsph<PERSIST>, blocks + nd, nbr_block, nd ? lds * npx_dense_lds_floats<6>() : 0,
                       sc, bs, cm, cst, cel, (const float4*)xs[cur]->p, (const float4*)VT.p, OM.p, VX.p, R, rd));
    NPX_TRY(launch_shm("vort_commit", k_vort_commit<PERSIST>, blocks + nd, nbr_block, nd ? lds * npx_dense_lds_floats<3>() : 0,
                       sc, bs, cm, cst, cel, (const float4*)xs[cur]->p, (const float4*)VT.p, (const float4*)OM.p,
                       (const float4*)VX.p, X.p, V.p, R, rd));
    return NPX_OK;
  }

  // Enqueues one step (no host synchronisation, so it can be captured).
  int enqueue_step() {
    int imax = 0;
    bool carry_rho = false;
    for (const auto& H : scenes) {
      if (!H.used || H.n == 0) continue;
      imax = std::max(imax, H.prm.solver_iterations);
      carry_rho = carry_rho || H.prm.solver_iterations <= 0;
    }
    NPX_TRY(grid_pass(carry_rho));
    swap_committed(carry_rho);
    const int nvb = cdiv((long long)hw_blocks * NPX_BLOCK, nbr_block);  // particle blocks of a neighbour pass
    // per-lane launch blocks: one per particle block, or the persistent count
    // (a multiple of 8) when that is smaller; each block then loops
    const int persist = (persist_blocks > 0 && persist_blocks < nvb) ? 1 : 0;
    const int blocks = persist ? persist_blocks : nvb;
    const Skin R{SKL.p, SKN.p, EXL.p, EXN.p, CDB.p, (unsigned*)CMAX.p, RSTART.p, PI.p, PS.p,
                 (const unsigned*)CMAX.p + total_cls(), CMAX.p + 2 * total_cls(), DL.p, DN.p, fast ? 0 : 1, LST.p,
                 (!fast && list_policy == 1) ? 1 : 0};
    // The build pass's dense list: one 64-lane block per batch, at most this
    // many at a time.  The later passes sweep the list in their first nd
    // launch blocks (a multiple of 8, for the XCD order of the rest), one
    // wave per batch, beside their per-lane blocks.
    const int dense_blocks = std::max(1, std::min(NPX_DENSE_WAVES, hw_blocks * (NPX_BLOCK / 64)));
    const int wpb = nbr_block / 64;
    const int nd = fast ? 0 : cdiv(std::min(NPX_DENSE_WAVES, dense_blocks), wpb * 8) * 8;
    const size_t lds = (size_t)wpb * sizeof(float);
    if (!fast) NPX_TRY(launch("dense_reset", k_dense_reset, 1, 256, DN.p, LST.p, max_scenes, list_policy));
    const BlockPlan rm{xcd_remap, 0, nvb, (const int*)XORD.p};
    const BlockPlan rd{xcd_remap, nd, nvb, (const int*)XORD.p};
    int cur = 0;
    const int npass = std::max(imax, 1);
    NPX_TRY(persist ? enqueue_passes<true>(npass, blocks, nd, dense_blocks, lds, R, rm, rd, cur)
                    : enqueue_passes<false>(npass, blocks, nd, dense_blocks, lds, R, rm, rd, cur));
    NPX_TRY(launch("run_fill", k_run_fill, hw_blocks, NPX_BLOCK, (const SceneC*)d_scenes.p,
                   (const int*)d_blk_scene.p, R, X.p, V.p, RHO.p, LAM.p));
    xs_final = cur;
    return NPX_OK;
  }

  // ---------------------------------------------------------- hipGraph
  // A step is captured once per parity of the committed-state buffers (the
  // step reads X and writes X2, then the roles swap), so two graphs cover
  // every step; they are re-captured when the launch shape or any buffer
  // changes.  Scene constants live in device memory, so dt / gravity / mass
  // updates need no re-capture.
  struct GraphSlot {
    hipGraphExec_t exec = nullptr;
    const void* in = nullptr;
  };
  GraphSlot graphs[2];
  std::vector<const void*> graph_key;
  int graph_captures = 0;

  // Position of the (vorticity * 2 + carry) entry in the graph key: a
  // replay swaps density and lambda exactly when the captured step did.
  static constexpr int kKeyCarry = 5;

  std::vector<const void*> current_graph_key() const {
    int imax = 0, vort = 0, carry = 0;
    for (const auto& H : scenes) {
      if (!H.used || H.n == 0) continue;
      imax = std::max(imax, H.prm.solver_iterations);
      vort |= H.prm.vorticity_epsilon != 0.0f;
      carry |= H.prm.solver_iterations <= 0;
    }
    // pointer pairs are order-insensitive (they swap every step)
    auto lo = [](const void* a, const void* b) { return a < b ? a : b; };
    auto hi = [](const void* a, const void* b) { return a < b ? b : a; };
    return {(const void*)(intptr_t)hw_blocks, (const void*)(intptr_t)nbr_block, (const void*)(intptr_t)xcd_remap, (const void*)(intptr_t)total_cls(),
            (const void*)(intptr_t)imax,
            (const void*)(intptr_t)(vort * 2 + carry), lo(X.p, X2.p), hi(X.p, X2.p), lo(V.p, V2.p),
            hi(V.p, V2.p), lo(ID.p, ID2.p), hi(ID.p, ID2.p), lo(CELL.p, CELL2.p), hi(CELL.p, CELL2.p),
            lo(RHO.p, RHO2.p), hi(RHO.p, RHO2.p), lo(LAM.p, LAM2.p), hi(LAM.p, LAM2.p), XS0.p, XS1.p, VT.p, OM.p, VX.p, clsmap.p, cstart.p, bsum.p,
            d_scenes.p, d_blk_scene.p, CLS.p, OFF.p, TPERM.p, TKEY.p, PERM.p, MPERM.p, MKEY.p, BIGB.p, HMARK.p, HIDP.p, HCNT.p, HUGE.p, XORD.p, SKL.p, SKN.p, EXL.p, EXN.p, CDB.p, CMAX.p, RSTART.p, PI.p, PS.p, BOFF.p, DL.p, DN.p, LST.p,
            (const void*)(intptr_t)persist_blocks, (const void*)(intptr_t)huge_levels,
            (const void*)(intptr_t)list_policy};
  }

  void drop_graphs() {
    bool any = false;
    for (const auto& g : graphs) any = any || g.exec != nullptr;
    if (any && stream) (void)hipStreamSynchronize(stream);  // a launch of an executable must end before it goes
    for (auto& g : graphs) {
      if (g.exec) (void)hipGraphExecDestroy(g.exec);
      g = GraphSlot();
    }
    graph_key.clear();
  }

  void swap_committed(bool carry_rho) {
    std::swap(X, X2);
    std::swap(V, V2);
    std::swap(ID, ID2);
    std::swap(CELL, CELL2);
    if (carry_rho) {
      std::swap(RHO, RHO2);
      std::swap(LAM, LAM2);
    }
  }

  int step_graph() {
    auto key = current_graph_key();
    if (key != graph_key) {
      drop_graphs();
      graph_key = key;
    }
    const int slot = (graphs[0].in == X.p) ? 0 : (graphs[1].in == X.p) ? 1 : -1;
    if (slot >= 0) {
      NPX_HIP(hipGraphLaunch(graphs[slot].exec, stream));
      swap_committed(((intptr_t)key[kKeyCarry] & 1) != 0);
      xs_final = graph_xs_final;
      return NPX_OK;
    }
    const int free_slot = graphs[0].exec ? 1 : 0;
    const void* in = X.p;
    hipGraph_t g = nullptr;
    NPX_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    int r = enqueue_step();
    hipError_t e = hipStreamEndCapture(stream, &g);
    if (r != NPX_OK) {
      if (g) (void)hipGraphDestroy(g);
      return r;
    }
    if (e != hipSuccess) return fail(NPX_EDEVICE, "hipStreamEndCapture: %s", hipGetErrorString(e));
    hipGraphExec_t exec = nullptr;
    e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return fail(NPX_EDEVICE, "hipGraphInstantiate: %s", hipGetErrorString(e));
    graphs[free_slot].exec = exec;
    graphs[free_slot].in = in;
    graph_xs_final = xs_final;
    ++graph_captures;
    // the capture swapped the host pointers as the step would; now run it
    NPX_HIP(hipGraphLaunch(exec, stream));
    return NPX_OK;
  }
  int graph_xs_final = 0;

  // Every step is replayed as a hipGraph (one capture per buffer parity and
  // launch shape; a step is about 30 launches whatever the scene count), unless
  // the context was created with NPX_FLAG_NO_GRAPH or per-kernel timing is on
  // (it records an event pair around each launch).
  int step_n(int nsteps) {
    for (int k = 0; k < nsteps; ++k) {
      if (!any_particles()) return NPX_OK;  // n == 0: PBFSolver::step returns (pbf_solver.cpp:21)
      NPX_TRY(prepare_step());
      const bool use_graph = !(flags & NPX_FLAG_NO_GRAPH) && !timing;
      if (use_graph)
        NPX_TRY(step_graph());
      else
        NPX_TRY(enqueue_step());
    }
    return NPX_OK;
  }

  // --------------------------------------------------------- diagnostics
  int debug_grid(uint64_t* keys, int32_t* cand, int32_t* nbr) {
    SceneH& H = scenes[0];
    if (!H.used || H.n == 0) return NPX_OK;
    NPX_TRY(prepare_step());
    NPX_TRY(grid_pass(false));
    const int n = H.n;
    DBuf<int> dc, dn;
    DBuf<unsigned long long> dk;
    NPX_HIP(dc.ensure(slot_cap));
    NPX_HIP(dn.ensure(slot_cap));
    NPX_HIP(dk.ensure(slot_cap));
    NPX_TRY(launch("debug_counts", k_debug_counts, hw_blocks, NPX_BLOCK, (const SceneC*)d_scenes.p,
                   (const int*)d_blk_scene.p, (const int*)clsmap.p, (const int*)cstart.p, (const int*)CELL2.p,
                   (const float4*)XS0.p, dc.p, dn.p, dk.p));
    std::vector<int> hc(n), hn(n), hid(n);
    std::vector<unsigned long long> hk(n);
    NPX_HIP(hipMemcpyAsync(hc.data(), dc.p, sizeof(int) * n, hipMemcpyDeviceToHost, stream));
    NPX_HIP(hipMemcpyAsync(hn.data(), dn.p, sizeof(int) * n, hipMemcpyDeviceToHost, stream));
    NPX_HIP(hipMemcpyAsync(hk.data(), dk.p, sizeof(unsigned long long) * n, hipMemcpyDeviceToHost, stream));
    NPX_HIP(hipMemcpyAsync(hid.data(), ID2.p, sizeof(int) * n, hipMemcpyDeviceToHost, stream));
    NPX_HIP(hipStreamSynchronize(stream));
    dc.release();
    dn.release();
    dk.release();
    NPX_TRY(check_flags());
    for (int q = 0; q < n; ++q) {
      const int i = hid[q];
      if (keys) keys[i] = hk[q];
      if (cand) cand[i] = hc[q];
      if (nbr) nbr[i] = hn[q];
    }
    return NPX_OK;
  }

  int debug_candidates(int32_t* lists, int64_t cap, int64_t* total) {
    SceneH& H = scenes[0];
    *total = 0;
    if (!H.used || H.n == 0) return NPX_OK;
    const int n = H.n;
    std::vector<int32_t> cnt(n);
    std::vector<uint64_t> keys(n);
    NPX_TRY(debug_grid(keys.data(), cnt.data(), nullptr));  // grid state stays in the *2 buffers
    std::vector<long long> off_id(n + 1, 0);
    for (int i = 0; i < n; ++i) off_id[i + 1] = off_id[i] + cnt[i];
    *total = off_id[n];
    if (!lists || cap < *total) return NPX_OK;
    std::vector<int> hid(n);
    NPX_TRY(copy_sync(hid.data(), ID2.p, sizeof(int) * n, hipMemcpyDeviceToHost));
    std::vector<long long> off_slot(n);
    for (int q = 0; q < n; ++q) off_slot[q] = off_id[hid[q]];
    DBuf<long long> doff;
    DBuf<int> dout;
    NPX_HIP(doff.ensure(n));
    NPX_HIP(dout.ensure((size_t)std::max<long long>(*total, 1)));
    NPX_TRY(copy_sync(doff.p, off_slot.data(), sizeof(long long) * n, hipMemcpyHostToDevice));
    NPX_TRY(launch("debug_lists", k_debug_lists, hw_blocks, NPX_BLOCK, (const SceneC*)d_scenes.p,
                   (const int*)d_blk_scene.p, (const int*)clsmap.p, (const int*)cstart.p, (const int*)CELL2.p,
                   (const int*)ID2.p, (const long long*)doff.p, dout.p));
    NPX_HIP(hipMemcpyAsync(lists, dout.p, sizeof(int) * (*total), hipMemcpyDeviceToHost, stream));
    NPX_HIP(hipStreamSynchronize(stream));
    doff.release();
    dout.release();
    return NPX_OK;
  }

  // --------------------------------------------------------- aux grids
  // Sorts n float4 points (payload in w) into the classes of window W
  // (clsmap values are class ids relative to cls_base), ordering each class
  // by key (NULL = point index; keys are unique in [0, n)).  Results:
  // Acstart (ncls+1), APERM, PS.
  int aux_sort(const GridWin& W, const int* cm, int cls_base, const float4* P, int n, const int* key, float4* PS) {
    NPX_HIP(Acstart.ensure((size_t)W.ncell + 1));
    NPX_HIP(ACLS.ensure(std::max(n, 1)));
    NPX_HIP(AOFF.ensure(std::max(n, 1)));
    NPX_HIP(ATPERM.ensure(std::max(n, 1)));
    NPX_HIP(ATKEY.ensure(std::max(n, 1)));
    NPX_HIP(APERM.ensure(std::max(n, 1)));
    NPX_HIP(AMPERM.ensure(std::max(n, 1)));
    NPX_HIP(AMKEY.ensure(std::max(n, 1)));
    NPX_HIP(ABIGB.ensure((size_t)cdiv(std::max(n, 1), NPX_BLOCK) + 1));
    int* cst = Acstart.p - cls_base;  // indexed by class id
    NPX_TRY(launch("aux_fill", k_fill_i32, cdiv(W.ncell + 1, 256), 256, Acstart.p, 0, W.ncell + 1));
    NPX_TRY(launch("aux_bin", k_aux_bin, cdiv(n, NPX_BIN_BLOCK), NPX_BIN_BLOCK, W, P, n, cm, cst, ACLS.p, AOFF.p));
    NPX_TRY(scan(Acstart.p, W.ncell + 1));
    HugeBufs hb;
    NPX_TRY(huge_prepare(AHID, AHIDP, AHCNT, AHUGE, AHMARK, W.ncell, n, cls_base, hb));
    NPX_TRY(launch("aux_scatter", k_aux_scatter, cdiv(n, 256), 256, n, (const int*)cst, (const int*)ACLS.p,
                   (const int*)AOFF.p, key, ATPERM.p, ATKEY.p, hb.hid, hb.huge, huge_levels));
    return aux_rank_all("aux_rank", n, W.ncell + cls_base, cst, ACLS.p, AOFF.p, ATPERM.p, ATKEY.p, AMPERM.p,
                        AMKEY.p, ABIGB.p, P, APERM.p, PS, &hb);
  }

  GridWin scene_win(const SceneH& H) const {
    GridWin W;
    W.inv_h = 1.0f / H.prm.kernel_radius;
    W.wlo0 = H.wlo[0];
    W.wlo1 = H.wlo[1];
    W.wlo2 = H.wlo[2];
    W.wx = H.wdim[0];
    W.wy = H.wdim[1];
    W.wz = H.wdim[2];
    W.ncell = H.ncell;
    return W;
  }
};

}  // namespace npx
