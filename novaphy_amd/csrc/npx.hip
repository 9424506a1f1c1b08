// npx.hip -- libnpx_hip: the MI355X PBF fluid path behind a C ABI.
//
// Build: hipcc --offload-arch=gfx950 -O3 -ffp-contract=off
//        -fhip-fp32-correctly-rounded-divide-sqrt -shared -fPIC
// (novaphy_amd/build.py).  The float operations of the default kernels are
// the reference's, one IEEE operation at a time, so results are bitwise equal
// to the FP32 CPU restatement of src/fluid/*.cpp on the same input.
#include <mutex>

#include "npx_engine.hpp"

using npx::Ctx;
using npx::SceneH;
using npx::cdiv;

namespace {

thread_local std::string g_create_error = "no error";

int check_params(Ctx* c, const npx_pbf_params* p) {
  if (!p) return c->fail(NPX_EINVAL, "params is NULL");
  if (!(p->kernel_radius > 0.0f) || !std::isfinite(p->kernel_radius))
    return c->fail(NPX_EINVAL, "kernel_radius must be positive and finite");
  if (!(p->rest_density > 0.0f)) return c->fail(NPX_EINVAL, "rest_density must be positive");
  if (p->solver_iterations < 0) return c->fail(NPX_EINVAL, "solver_iterations must be >= 0");
  return NPX_OK;
}

// ---------------------------------------------------------------- boundary
int set_boundary(Ctx* c, const float* local3, const int32_t* body, const float* psi, int nb, int nbodies) {
  if (nb < 0 || nbodies < 0) return c->fail(NPX_EINVAL, "negative boundary size");
  if (nb > 0 && (!local3 || !body)) return c->fail(NPX_EINVAL, "boundary arrays are NULL");
  c->nb = nb;
  c->nbodies = nbodies;
  const size_t n = std::max(nb, 1);
  NPX_HIP_C(c, c->BL.ensure(n));
  NPX_HIP_C(c, c->BP.ensure(n));
  NPX_HIP_C(c, c->BPS.ensure(n));
  NPX_HIP_C(c, c->FB.ensure(n));
  NPX_HIP_C(c, c->BODY.ensure(n));
  NPX_HIP_C(c, c->PSI.ensure(n));
  NPX_HIP_C(c, c->BODY_PTR.ensure((size_t)nbodies + 1));
  NPX_HIP_C(c, c->BODY_IDX.ensure(n));
  NPX_HIP_C(c, c->POSE.ensure((size_t)std::max(nbodies, 1) * 7));
  NPX_HIP_C(c, c->STATIC.ensure((size_t)std::max(nbodies, 1)));
  NPX_HIP_C(c, c->OUT6.ensure((size_t)std::max(nbodies, 1) * 6));
  std::vector<float4> bl(n);
  std::vector<float> ps(n, 0.0f);
  std::vector<int> bd(n, -1);
  std::vector<int> cnt(nbodies + 1, 0);
  for (int b = 0; b < nb; ++b) {
    bl[b] = make_float4(local3[3 * b], local3[3 * b + 1], local3[3 * b + 2], 0.0f);
    bd[b] = body[b];
    if (psi) ps[b] = psi[b];
    if (body[b] >= 0 && body[b] < nbodies) cnt[body[b] + 1]++;
  }
  for (int k = 0; k < nbodies; ++k) cnt[k + 1] += cnt[k];
  std::vector<int> idx(n, 0), fillp(cnt.begin(), cnt.end());
  for (int b = 0; b < nb; ++b)
    if (body[b] >= 0 && body[b] < nbodies) idx[fillp[body[b]]++] = b;  // ascending b per body
  NPX_TRY(c->copy_sync(c->BL.p, bl.data(), sizeof(float4) * n, hipMemcpyHostToDevice));
  NPX_TRY(c->copy_sync(c->BODY.p, bd.data(), sizeof(int) * n, hipMemcpyHostToDevice));
  NPX_TRY(c->copy_sync(c->PSI.p, ps.data(), sizeof(float) * n, hipMemcpyHostToDevice));
  NPX_TRY(c->copy_sync(c->BODY_PTR.p, cnt.data(), sizeof(int) * (nbodies + 1), hipMemcpyHostToDevice));
  NPX_TRY(c->copy_sync(c->BODY_IDX.p, idx.data(), sizeof(int) * n, hipMemcpyHostToDevice));
  return NPX_OK;
}

int upload_poses(Ctx* c, const float* pose7, int nbodies) {
  if (nbodies != c->nbodies) return c->fail(NPX_EINVAL, "expected %d body poses, got %d", c->nbodies, nbodies);
  if (nbodies > 0) {
    if (!pose7) return c->fail(NPX_EINVAL, "pose7 is NULL");
    NPX_TRY(c->upload_bodies(c->POSE.p, pose7, nullptr, nullptr, nbodies));
  }
  return NPX_OK;
}

int bpos(Ctx* c, const float* pose7, int nbodies) {
  NPX_TRY(upload_poses(c, pose7, nbodies));
  NPX_TRY(c->launch("bpos", k_bpos, cdiv(c->nb, 256), 256, (const float4*)c->BL.p, (const int*)c->BODY.p,
                    (const float*)c->PSI.p, c->nb, (const float*)c->POSE.p, nbodies, c->BP.p));
  return NPX_OK;
}

int compute_volumes(Ctx* c, const float* pose7, int nbodies, float* psi_out) {
  if (c->nb == 0) return NPX_OK;
  NPX_TRY(bpos(c, pose7, nbodies));
  // window = bounding box of the boundary particles (+2 cells)
  std::vector<float4> bp(c->nb);
  NPX_HIP_C(c, hipMemcpyAsync(bp.data(), c->BP.p, sizeof(float4) * c->nb, hipMemcpyDeviceToHost, c->stream));
  NPX_HIP_C(c, hipStreamSynchronize(c->stream));
  const float h = c->scenes[0].prm.kernel_radius;
  const float inv_h = 1.0f / h;
  int lo[3] = {0x7FFFFFFF, 0x7FFFFFFF, 0x7FFFFFFF}, hi[3] = {-0x7FFFFFFF, -0x7FFFFFFF, -0x7FFFFFFF};
  for (int b = 0; b < c->nb; ++b) {
    const float p[3] = {bp[b].x, bp[b].y, bp[b].z};
    for (int d = 0; d < 3; ++d) {
      const float f = std::floor(p[d] * inv_h);
      if (!(std::fabs(f) < 1.0e8f)) return c->fail(NPX_EDOMAIN, "boundary particle position is not finite");
      lo[d] = std::min(lo[d], (int)f);
      hi[d] = std::max(hi[d], (int)f);
    }
  }
  int wlo[3], wdim[3];
  long long ncell = 1;
  for (int d = 0; d < 3; ++d) {
    wlo[d] = lo[d] - 2;
    wdim[d] = hi[d] + 2 - wlo[d] + 1;
    ncell *= wdim[d];
  }
  if (ncell > (1ll << 30)) return c->fail(NPX_EDOMAIN, "boundary extent too large for a cell window");
  NPX_HIP_C(c, c->aux_clsmap.ensure((size_t)ncell));
  NPX_TRY(c->build_class_table(wlo, wdim, 0, c->aux_clsmap.p));
  GridWin W;
  W.inv_h = inv_h;
  W.wlo0 = wlo[0];
  W.wlo1 = wlo[1];
  W.wlo2 = wlo[2];
  W.wx = wdim[0];
  W.wy = wdim[1];
  W.wz = wdim[2];
  W.ncell = (int)ncell;
  NPX_TRY(c->aux_sort(W, c->aux_clsmap.p, 0, c->BP.p, c->nb, nullptr, c->BPS.p));
  const float h2 = h * h;
  const float h9 = h2 * h2 * h2 * h2 * h;
  const float c6 = 315.0f / ((64.0f * npx::kPI) * h9);
  NPX_TRY(c->launch("psi", k_psi, cdiv(c->nb, 256), 256, W, h2, c6, (const float4*)c->BP.p, c->nb,
                    (const int*)c->aux_clsmap.p, (const int*)c->Acstart.p, (const float4*)c->BPS.p, c->PSI.p));
  if (psi_out)
    NPX_HIP_C(c, hipMemcpyAsync(psi_out, c->PSI.p, sizeof(float) * c->nb, hipMemcpyDeviceToHost, c->stream));
  NPX_HIP_C(c, hipStreamSynchronize(c->stream));
  return NPX_OK;
}

int boundary_density(Ctx* c) {
  SceneH& H = c->scenes[0];
  if (c->nb == 0 || !H.used || H.n == 0) return NPX_OK;
  if (!H.window_valid) return c->fail(NPX_ESTATE, "no step has run yet");
  const GridWin W = c->scene_win(H);
  const int* cm = c->clsmap.p + H.cls_base;
  NPX_TRY(c->aux_sort(W, cm, H.cls_base, c->BP.p, c->nb, nullptr, c->BPS.p));
  NPX_TRY(c->launch("boundary_density", k_boundary_density, c->hw_blocks, NPX_BLOCK, (const SceneC*)c->d_scenes.p,
                    (const int*)c->d_blk_scene.p, W, (const float4*)c->X.p, cm,
                    (const int*)(c->Acstart.p - H.cls_base), (const float4*)c->BPS.p, c->RHO.p, c->d_flags.p));
  return NPX_OK;
}

int coupling(Ctx* c, const float* pose7, const uint8_t* is_static, int nbodies, float* force3, float* torque3) {
  SceneH& H = c->scenes[0];
  if (nbodies != c->nbodies) return c->fail(NPX_EINVAL, "expected %d bodies, got %d", c->nbodies, nbodies);
  if (nbodies == 0) return NPX_OK;
  std::vector<float> out(6 * nbodies, 0.0f);
  if (c->nb > 0 && H.used && H.n > 0) {
    if (!H.window_valid) return c->fail(NPX_ESTATE, "no step has run yet");
    if (!pose7 || !is_static) return c->fail(NPX_EINVAL, "pose7 / is_static is NULL");
    NPX_TRY(c->upload_bodies(c->POSE.p, pose7, c->STATIC.p, is_static, nbodies));
    NPX_HIP_C(c, c->FP.ensure(c->slot_cap));
    NPX_HIP_C(c, c->FS.ensure(c->slot_cap));
    NPX_HIP_C(c, c->BFK.ensure(c->slot_cap));
    NPX_TRY(c->launch("fluid_payload", k_fluid_payload, c->hw_blocks, NPX_BLOCK, (const SceneC*)c->d_scenes.p,
                      (const int*)c->d_blk_scene.p, (const float4*)c->X.p, (const float*)c->RHO.p, c->FP.p,
                      (const int*)c->ID.p, c->BFK.p));
    const GridWin W = c->scene_win(H);
    const int* cm = c->clsmap.p + H.cls_base;
    NPX_TRY(c->aux_sort(W, cm, H.cls_base, c->FP.p + H.seg_off, H.n, c->ID.p + H.seg_off, c->FS.p));
    NPX_TRY(c->payload_runs(H.n, (const int*)(c->Acstart.p - H.cls_base), W.ncell + H.cls_base,
                            (const int*)c->AOFF.p, (const float4*)c->FS.p));
    const float h = H.prm.kernel_radius;
    const float h6 = h * h * h * h * h * h;
    const float cs = 45.0f / (npx::kPI * h6);
    // a wave per boundary particle
    NPX_TRY(c->launch("coupling", k_coupling_particles, cdiv(c->nb, 4), 256, W, h, h * h, cs, H.prm.rest_density,
                      H.mass, (const float4*)c->BP.p, (const int*)c->BODY.p, c->nb, (const uint8_t*)c->STATIC.p,
                      nbodies, cm, (const int*)(c->Acstart.p - H.cls_base), (const float4*)c->FS.p,
                      (const int*)c->CRPI.p, (const int*)c->CRPS.p, c->FB.p));
    NPX_TRY(c->launch("body_reduce", k_body_reduce, nbodies, 64, (const int*)c->BODY_PTR.p,
                      (const int*)c->BODY_PTR.p + 1, (const int*)c->BODY_IDX.p, nbodies, (const float4*)c->FB.p, (const float4*)c->BP.p,
                      (const float*)c->POSE.p, c->OUT6.p));
    NPX_TRY(c->read_bodies(out.data(), c->OUT6.p, nbodies));
    NPX_TRY(c->check_flags());
  }
  for (int k = 0; k < nbodies; ++k) {
    if (force3) {
      force3[3 * k] = out[6 * k];
      force3[3 * k + 1] = out[6 * k + 1];
      force3[3 * k + 2] = out[6 * k + 2];
    }
    if (torque3) {
      torque3[3 * k] = out[6 * k + 3];
      torque3[3 * k + 1] = out[6 * k + 4];
      torque3[3 * k + 2] = out[6 * k + 5];
    }
  }
  return NPX_OK;
}

// ------------------------------------------------------- standalone grid
int grid_query(Ctx* c, const float* pos3, int n, float cell, const float* q3, int nq, int32_t* counts,
               int32_t* lists, int64_t cap, int64_t* total) {
  *total = 0;
  if (nq == 0) return NPX_OK;
  if (!(cell > 0.0f)) return c->fail(NPX_EINVAL, "cell_size must be positive");
  const float inv = 1.0f / cell;
  int lo[3] = {0x7FFFFFFF, 0x7FFFFFFF, 0x7FFFFFFF}, hi[3] = {-0x7FFFFFFF, -0x7FFFFFFF, -0x7FFFFFFF};
  auto acc = [&](const float* p) -> bool {
    for (int d = 0; d < 3; ++d) {
      const float f = std::floor(p[d] * inv);
      if (!(std::fabs(f) < 1.0e8f)) return false;
      lo[d] = std::min(lo[d], (int)f);
      hi[d] = std::max(hi[d], (int)f);
    }
    return true;
  };
  for (int i = 0; i < n; ++i)
    if (!acc(pos3 + 3 * i)) return c->fail(NPX_EDOMAIN, "point %d is not finite", i);
  for (int i = 0; i < nq; ++i)
    if (!acc(q3 + 3 * i)) return c->fail(NPX_EDOMAIN, "query %d is not finite", i);
  int wlo[3], wdim[3];
  long long ncell = 1;
  for (int d = 0; d < 3; ++d) {
    wlo[d] = lo[d] - 2;
    wdim[d] = hi[d] + 2 - wlo[d] + 1;
    ncell *= wdim[d];
  }
  if (ncell > (1ll << 28)) return c->fail(NPX_EDOMAIN, "points span too many cells for one window (%lld)", ncell);
  NPX_HIP_C(c, c->aux_clsmap.ensure((size_t)ncell));
  NPX_TRY(c->build_class_table(wlo, wdim, 0, c->aux_clsmap.p));
  GridWin W;
  W.inv_h = inv;
  W.wlo0 = wlo[0];
  W.wlo1 = wlo[1];
  W.wlo2 = wlo[2];
  W.wx = wdim[0];
  W.wy = wdim[1];
  W.wz = wdim[2];
  W.ncell = (int)ncell;
  npx::DBuf<float4> P, PS, Q;
  npx::DBuf<int> cnt;
  npx::DBuf<long long> off;
  npx::DBuf<int> out;
  NPX_HIP_C(c, P.ensure(std::max(n, 1)));
  NPX_HIP_C(c, PS.ensure(std::max(n, 1)));
  NPX_HIP_C(c, Q.ensure(nq));
  NPX_HIP_C(c, cnt.ensure(nq));
  NPX_HIP_C(c, off.ensure(nq));
  std::vector<float4> hp(std::max(n, 1)), hq(nq);
  for (int i = 0; i < n; ++i) hp[i] = make_float4(pos3[3 * i], pos3[3 * i + 1], pos3[3 * i + 2], 0.0f);
  for (int i = 0; i < nq; ++i) hq[i] = make_float4(q3[3 * i], q3[3 * i + 1], q3[3 * i + 2], 0.0f);
  if (n > 0) NPX_TRY(c->copy_sync(P.p, hp.data(), sizeof(float4) * n, hipMemcpyHostToDevice));
  NPX_TRY(c->copy_sync(Q.p, hq.data(), sizeof(float4) * nq, hipMemcpyHostToDevice));
  NPX_TRY(c->aux_sort(W, c->aux_clsmap.p, 0, P.p, n, nullptr, PS.p));
  NPX_TRY(c->launch("query_count", k_query_count, cdiv(nq, 256), 256, W, (const float4*)Q.p, nq,
                    (const int*)c->aux_clsmap.p, (const int*)c->Acstart.p, cnt.p));
  std::vector<int> hc(nq);
  NPX_TRY(c->copy_sync(hc.data(), cnt.p, sizeof(int) * nq, hipMemcpyDeviceToHost));
  std::vector<long long> ho(nq);
  long long t = 0;
  for (int i = 0; i < nq; ++i) {
    ho[i] = t;
    t += hc[i];
    if (counts) counts[i] = hc[i];
  }
  *total = t;
  if (lists && cap >= t && t > 0) {
    NPX_HIP_C(c, out.ensure((size_t)t));
    NPX_TRY(c->copy_sync(off.p, ho.data(), sizeof(long long) * nq, hipMemcpyHostToDevice));
    NPX_TRY(c->launch("query_lists", k_query_lists, cdiv(nq, 256), 256, W, (const float4*)Q.p, nq,
                      (const int*)c->aux_clsmap.p, (const int*)c->Acstart.p, (const int*)c->APERM.p,
                      (const long long*)off.p, out.p));
    NPX_TRY(c->copy_sync(lists, out.p, sizeof(int) * t, hipMemcpyDeviceToHost));
  }
  P.release();
  PS.release();
  Q.release();
  cnt.release();
  off.release();
  out.release();
  return NPX_OK;
}

void destroy(Ctx* c) {
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  c->drop_graphs();
  for (auto& t : c->timers) {
    (void)hipEventDestroy(t.a);
    (void)hipEventDestroy(t.b);
  }
  for (auto e : c->ev_pool) (void)hipEventDestroy(e);
  npx::DBuf<float4>* f4[] = {&c->X, &c->X2, &c->V, &c->V2, &c->XS0, &c->XS1, &c->VT, &c->OM, &c->VX,
                             &c->BL, &c->BP, &c->BPS, &c->FP, &c->FS, &c->FB};
  for (auto* b : f4) b->release();
  npx::DBuf<float>* f1[] = {&c->RHO, &c->RHO2, &c->LAM, &c->LAM2, &c->CDB, &c->PSI, &c->POSE, &c->OUT6};
  for (auto* b : f1) b->release();
  npx::DBuf<int>* i1[] = {&c->ID, &c->ID2, &c->CELL, &c->CELL2, &c->CLS, &c->OFF, &c->TPERM, &c->TKEY,
                          &c->PERM, &c->MPERM, &c->MKEY, &c->HIDP, &c->HCNT, &c->HUGE, &c->XORD, &c->SKL, &c->SKN, &c->EXL, &c->EXN, &c->CMAX, &c->clsmap, &c->cstart, &c->bsum, &c->htv, &c->d_blk_scene, &c->d_flags,
                          &c->BODY, &c->BODY_PTR, &c->BODY_IDX, &c->Acstart, &c->ACLS, &c->AOFF, &c->ATPERM,
                          &c->ATKEY, &c->APERM, &c->AMPERM, &c->AMKEY, &c->aux_clsmap, &c->CRF, &c->CRPI,
                          &c->CRPS};
  for (auto* b : i1) b->release();
  c->htk.release();
  c->BIGB.release();
  c->HMARK.release();
  c->ABIGB.release();
  c->RSTART.release();
  c->PI.release();
  c->PS.release();
  c->BOFF.release();
  c->DL.release();
  c->DN.release();
  c->LST.release();
  for (auto* ps : {&c->bsort, &c->fsort}) {
    ps->cst.release();
    ps->cls.release();
    ps->off.release();
    ps->tperm.release();
    ps->tkey.release();
    ps->perm.release();
    ps->mperm.release();
    ps->mkey.release();
    ps->bigb.release();
    ps->hid.release();
    ps->hidp.release();
    ps->hcnt.release();
    ps->huge.release();
    ps->hmark.release();
  }
  for (auto* b : {&c->BBL, &c->BBP, &c->BBPS, &c->BFP, &c->BFS, &c->BFB}) b->release();
  for (auto* b : {&c->BBODY, &c->BSCN, &c->BBIDX, &c->FPSC, &c->BFK, &c->BBSTART, &c->BBEND, &c->BSBASE}) b->release();
  for (auto* b : {&c->AHID, &c->AHIDP, &c->AHCNT, &c->AHUGE}) b->release();
  c->AHMARK.release();
  for (auto* b : {&c->BPSI, &c->BPOSE, &c->BOUT6}) b->release();
  c->BSTAT.release();
  c->d_scenes.release();
  c->STATIC.release();
  if (c->h_flags) (void)hipHostFree(c->h_flags);
  if (c->h_pose) (void)hipHostFree(c->h_pose);
  if (c->h_static) (void)hipHostFree(c->h_static);
  if (c->h_out) (void)hipHostFree(c->h_out);
  if (c->ev_up) (void)hipEventDestroy(c->ev_up);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

// ------------------------------------------------ allocation-free grid query
// npx_grid_query_noalloc keeps every device buffer it needs in `scratch`, a
// file-scope device array: each GPU the library runs on holds its own copy
// (NPX_NA_SCRATCH_BYTES, about 23.0 MiB) from module load until the process
// ends, and the call allocates nothing.  Host code reaches a GPU's copy only
// through the address hipGetSymbolAddress returns for it on that device
// (na_scratch_ptr); the array's name is not a device address on the host.
// Every call on a GPU shares that GPU's copy: calls that overlap on one GPU
// race.  What must start at zero is cleared by kernels on the call's stream.
#define NPX_NA_MAX_POINTS (1 << 16)
#define NPX_NA_MAX_QUERIES (1 << 16)
#define NPX_NA_MAX_CELLS (1 << 18)
#define NPX_NA_MAX_LIST (1 << 21)
#define NPX_NA_MAX_DEVICES 64

constexpr size_t al(size_t b) { return (b + 255) & ~(size_t)255; }

// Byte offsets of the regions (each sized for the maxima, 256-aligned).
struct NaLayout {
  static constexpr size_t HT = 2 * (size_t)NPX_NA_MAX_CELLS;  // hash-table capacity bound (power of two)
  static constexpr size_t htk = 0;
  static constexpr size_t htv = htk + al(HT * 8);
  static constexpr size_t clsmap = htv + al(HT * 4);
  static constexpr size_t cst = clsmap + al((size_t)NPX_NA_MAX_CELLS * 4);
  static constexpr size_t bsum = cst + al(((size_t)NPX_NA_MAX_CELLS + 1) * 4);
  static constexpr size_t raw_p = bsum + al(((size_t)NPX_NA_MAX_CELLS / 1024 + 2) * 4);
  static constexpr size_t P = raw_p + al((size_t)NPX_NA_MAX_POINTS * 12);
  static constexpr size_t PS = P + al((size_t)NPX_NA_MAX_POINTS * 16);
  static constexpr size_t raw_q = PS + al((size_t)NPX_NA_MAX_POINTS * 16);
  static constexpr size_t Q = raw_q + al((size_t)NPX_NA_MAX_QUERIES * 12);
  static constexpr size_t cls = Q + al((size_t)NPX_NA_MAX_QUERIES * 16);
  static constexpr size_t off = cls + al((size_t)NPX_NA_MAX_POINTS * 4);
  static constexpr size_t tperm = off + al((size_t)NPX_NA_MAX_POINTS * 4);
  static constexpr size_t tkey = tperm + al((size_t)NPX_NA_MAX_POINTS * 4);
  static constexpr size_t perm = tkey + al((size_t)NPX_NA_MAX_POINTS * 4);
  static constexpr size_t mperm = perm + al((size_t)NPX_NA_MAX_POINTS * 4);
  static constexpr size_t mkey = mperm + al((size_t)NPX_NA_MAX_POINTS * 4);
  static constexpr size_t bigb = mkey + al((size_t)NPX_NA_MAX_POINTS * 4);
  static constexpr size_t cnt = bigb + al((size_t)NPX_NA_MAX_POINTS / 256 + 2);
  static constexpr size_t off64 = cnt + al(((size_t)NPX_NA_MAX_QUERIES + 1) * 4);
  static constexpr size_t out = off64 + al((size_t)NPX_NA_MAX_QUERIES * 8);
  static constexpr size_t total = out + al((size_t)NPX_NA_MAX_LIST * 4);
};
#define NPX_NA_SCRATCH_BYTES (NaLayout::total)

static __device__ __align__(256) unsigned char scratch[NPX_NA_SCRATCH_BYTES];

std::mutex g_na_mu;
unsigned char* g_na_ptr[NPX_NA_MAX_DEVICES] = {};

// The current device's copy of `scratch`, looked up once per device with
// that device current.
int na_scratch_ptr(unsigned char** out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NPX_NA_MAX_DEVICES) {
    g_create_error = "npx_grid_query_noalloc: no usable HIP device";
    return NPX_EDEVICE;
  }
  std::lock_guard<std::mutex> lk(g_na_mu);
  if (!g_na_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      g_create_error = std::string("npx_grid_query_noalloc: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
      return NPX_EDEVICE;
    }
    g_na_ptr[dev] = static_cast<unsigned char*>(p);
  }
  *out = g_na_ptr[dev];
  return NPX_OK;
}

// Launch on the null stream; an empty grid is no launch.
template <class K, class... A>
hipError_t na_launch(K kern, int blocks, int threads, A... args) {
  if (blocks <= 0) return hipSuccess;
  hipLaunchKernelGGL(kern, dim3(blocks), dim3(threads), 0, 0, args...);
  return hipGetLastError();
}

// Every entry point of the C ABI runs its body inside guard(): no C++
// exception crosses the boundary.  std::bad_alloc (host allocation) becomes
// NPX_ENOMEM, any other exception NPX_EINTERNAL; the message goes to the
// context (or to the create-error slot when there is none).
void set_guard_error(npx_ctx* ctx, const char* fn, const char* what) noexcept {
  try {
    std::string m = std::string(fn) + ": " + what;
    if (ctx)
      reinterpret_cast<Ctx*>(ctx)->err = std::move(m);
    else
      g_create_error = std::move(m);
  } catch (...) {
  }
}

template <class F>
int guard(npx_ctx* ctx, const char* fn, F&& f) noexcept {
  try {
    return f();
  } catch (const std::bad_alloc&) {
    set_guard_error(ctx, fn, "host memory allocation failed");
    return NPX_ENOMEM;
  } catch (const std::exception& e) {
    set_guard_error(ctx, fn, e.what());
    return NPX_EINTERNAL;
  } catch (...) {
    set_guard_error(ctx, fn, "unknown exception");
    return NPX_EINTERNAL;
  }
}

}  // namespace

// ====================================================================== C ABI
extern "C" {

int32_t npx_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

const char* npx_last_error(const npx_ctx* ctx) {
  if (!ctx) return g_create_error.c_str();
  return reinterpret_cast<const Ctx*>(ctx)->err.c_str();
}

int npx_create(const npx_pbf_params* params, int32_t capacity, uint32_t flags, npx_ctx** out) {
  return guard(nullptr, __func__, [&]() -> int {
    if (!out) return NPX_EINVAL;
    *out = nullptr;
    if (capacity > NPX_MAX_SLOTS) {
      g_create_error = "npx_create: capacity beyond the slot limit (2^28)";
      return NPX_EINVAL;
    }
    Ctx* c = new (std::nothrow) Ctx();
    if (!c) {
      g_create_error = "host memory allocation failed";
      return NPX_ENOMEM;
    }
    int r = check_params(c, params);
    if (r == NPX_OK) r = c->init(std::max(capacity, 1), 0, 1, flags, false);
    if (r != NPX_OK) {
      g_create_error = c->err;
      destroy(c);
      return r;
    }
    c->scenes[0].prm = *params;
    *out = reinterpret_cast<npx_ctx*>(c);
    return NPX_OK;
  });
}

int npx_destroy(npx_ctx* ctx) {
  return guard(ctx, __func__, [&]() -> int {
    destroy(reinterpret_cast<Ctx*>(ctx));
    return NPX_OK;
  });
}

int npx_set_params(npx_ctx* ctx, const npx_pbf_params* params) {
  return guard(ctx, __func__, [&]() -> int {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c) return NPX_EINVAL;
    NPX_TRY(check_params(c, params));
    SceneH& H = c->scenes[0];
    const bool window_change = H.prm.kernel_radius != params->kernel_radius ||
                               H.prm.use_domain_bounds != params->use_domain_bounds ||
                               std::memcmp(H.prm.domain_lower, params->domain_lower, sizeof(float) * 3) != 0 ||
                               std::memcmp(H.prm.domain_upper, params->domain_upper, sizeof(float) * 3) != 0;
    H.prm = *params;
    c->table_dirty = true;
    if (window_change && H.used) {
      H.window_valid = false;
      NPX_TRY(c->scene_window(H));
    }
    return NPX_OK;
  });
}

int npx_set_particles(npx_ctx* ctx, const float* pos3, const float* vel3, int32_t n) {
  return guard(ctx, __func__, [&]() -> int {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c) return NPX_EINVAL;
    if (c->batch) return c->fail(NPX_ESTATE, "use npx_scene_add on a batch context");
    if (n < 0 || (n > 0 && !pos3)) return c->fail(NPX_EINVAL, "bad particle arrays");
    NPX_TRY(c->reserve_slots(n));
    SceneH& H = c->scenes[0];
    H.used = true;
    H.n = n;
    H.seg_off = 0;
    H.seg_blocks = cdiv(std::max(n, 1), NPX_BLOCK);
    std::fill(c->h_blk_scene.begin(), c->h_blk_scene.end(), -1);
    for (int b = 0; b < cdiv(n, NPX_BLOCK); ++b) c->h_blk_scene[b] = 0;
    c->table_dirty = true;
    H.window_valid = false;
    NPX_TRY(c->scene_window(H));
    NPX_TRY(c->upload_particles(H, pos3, vel3));
    NPX_TRY(c->upload_tables());
    NPX_TRY(c->reset_list_stats(0));
    return NPX_OK;
  });
}

int npx_get_particles(npx_ctx* ctx, float* pos3, float* vel3, float* rho, float* lambda) {
  return guard(ctx, __func__, [&]() -> int {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c) return NPX_EINVAL;
    return c->download(c->scenes[0], pos3, vel3, rho, lambda);
  });
}

int32_t npx_num_particles(const npx_ctx* ctx) {
  const Ctx* c = reinterpret_cast<const Ctx*>(ctx);
  if (!c) return 0;
  int t = 0;
  for (const auto& H : c->scenes)
    if (H.used) t += H.n;
  return t;
}

int npx_step_n(npx_ctx* ctx, float dt, const float gravity[3], float particle_mass, int32_t nsteps) {
  return guard(ctx, __func__, [&]() -> int {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c) return NPX_EINVAL;
    if (c->batch) return c->fail(NPX_ESTATE, "use npx_batch_step on a batch context");
    if (!gravity) return c->fail(NPX_EINVAL, "gravity is NULL");
    SceneH& H = c->scenes[0];
    if (H.dt != dt || H.mass != particle_mass || std::memcmp(H.g, gravity, sizeof(float) * 3) != 0) {
      H.dt = dt;
      H.mass = particle_mass;
      std::memcpy(H.g, gravity, sizeof(float) * 3);
      c->table_dirty = true;
    }
    if (!H.used || H.n == 0) return NPX_OK;
    return c->step_n(nsteps);
  });
}

int npx_step(npx_ctx* ctx, float dt, const float gravity[3], float particle_mass) {
  return guard(ctx, __func__, [&]() -> int {
    return npx_step_n(ctx, dt, gravity, particle_mass, 1);
  });
}

int npx_synchronize(npx_ctx* ctx) {
  return guard(ctx, __func__, [&]() -> int {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c) return NPX_EINVAL;
    NPX_HIP_C(c, hipStreamSynchronize(c->stream));
    return c->check_flags();
  });
}

int npx_debug_grid(npx_ctx* ctx, uint64_t* keys, int32_t* cand_cnt, int32_t* nbr_cnt) {
  return guard(ctx, __func__, [&]() -> int {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c) return NPX_EINVAL;
    if (c->batch) return c->fail(NPX_ESTATE, "grid diagnostics are for single-scene contexts");
    return c->debug_grid(keys, cand_cnt, nbr_cnt);
  });
}

int npx_debug_candidates(npx_ctx* ctx, int32_t* lists, int64_t cap, int64_t* total) {
  return guard(ctx, __func__, [&]() -> int {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c || !total) return NPX_EINVAL;
    if (c->batch) return c->fail(NPX_ESTATE, "grid diagnostics are for single-scene contexts");
    return c->debug_candidates(lists, cap, total);
  });
}

int npx_grid_query(const float* pos3, int32_t n, float cell_size, const float* query3, int32_t nq, int32_t* counts,
                   int32_t* lists, int64_t cap, int64_t* total) {
  return guard(nullptr, __func__, [&]() -> int {
    if (!total || n < 0 || nq < 0 || (n > 0 && !pos3) || (nq > 0 && !query3)) {
      g_create_error = "npx_grid_query: bad arguments";
      return NPX_EINVAL;
    }
    Ctx* c = new (std::nothrow) Ctx();
    if (!c) {
      g_create_error = "host memory allocation failed";
      return NPX_ENOMEM;
    }
    npx_pbf_params p{};
    p.kernel_radius = cell_size;
    p.rest_density = 1.0f;
    int r = c->init(1, 0, 1, 0, false);
    if (r == NPX_OK) {
      c->scenes[0].prm = p;
      r = grid_query(c, pos3, n, cell_size, query3, nq, counts, lists, cap, total);
    }
    if (r != NPX_OK) g_create_error = c->err;
    destroy(c);
    return r;
  });
}

int npx_grid_query_noalloc(const float* pos3, int32_t n, float cell_size, const float* query3, int32_t nq,
                           int32_t* counts, int32_t* lists, int64_t cap, int64_t* total) {
  return guard(nullptr, __func__, [&]() -> int {
    auto bad = [](int code, const char* msg) {
      g_create_error = std::string("npx_grid_query_noalloc: ") + msg;
      return code;
    };
    if (!total || n < 0 || nq < 0 || (n > 0 && !pos3) || (nq > 0 && !query3)) return bad(NPX_EINVAL, "bad arguments");
    *total = 0;
    if (n > NPX_NA_MAX_POINTS) return bad(NPX_EINVAL, "more than 65536 points");
    if (nq > NPX_NA_MAX_QUERIES) return bad(NPX_EINVAL, "more than 65536 queries");
    if (nq == 0) return NPX_OK;
    if (!(cell_size > 0.0f)) return bad(NPX_EINVAL, "cell_size must be positive");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return bad(NPX_EDEVICE, "no HIP device available");
    // window: bounding cells of points and queries, 2 cells of margin (as npx_grid_query)
    const float inv = 1.0f / cell_size;
    int lo[3] = {0x7FFFFFFF, 0x7FFFFFFF, 0x7FFFFFFF}, hi[3] = {-0x7FFFFFFF, -0x7FFFFFFF, -0x7FFFFFFF};
    for (int k = 0; k < n + nq; ++k) {
      const float* p = k < n ? pos3 + 3 * (size_t)k : query3 + 3 * (size_t)(k - n);
      for (int d = 0; d < 3; ++d) {
        const float f = std::floor(p[d] * inv);
        if (!(std::fabs(f) < 1.0e8f)) return bad(NPX_EDOMAIN, "a point or query is not finite");
        lo[d] = std::min(lo[d], (int)f);
        hi[d] = std::max(hi[d], (int)f);
      }
    }
    GridWin W;
    long long ncell = 1;
    int wdim[3], wlo[3];
    for (int d = 0; d < 3; ++d) {
      wlo[d] = lo[d] - 2;
      wdim[d] = hi[d] + 2 - wlo[d] + 1;
      ncell *= wdim[d];
    }
    if (ncell > NPX_NA_MAX_CELLS) return bad(NPX_EDOMAIN, "points span more than 2^18 cells");
    W.inv_h = inv;
    W.wlo0 = wlo[0];
    W.wlo1 = wlo[1];
    W.wlo2 = wlo[2];
    W.wx = wdim[0];
    W.wy = wdim[1];
    W.wz = wdim[2];
    W.ncell = (int)ncell;
    const int nc = (int)ncell;
    unsigned char* base = nullptr;
    NPX_TRY(na_scratch_ptr(&base));
    using L = NaLayout;
    auto* htk = reinterpret_cast<unsigned long long*>(base + L::htk);
    auto* htv = reinterpret_cast<int*>(base + L::htv);
    auto* clsmap = reinterpret_cast<int*>(base + L::clsmap);
    auto* cst = reinterpret_cast<int*>(base + L::cst);
    auto* bsum = reinterpret_cast<int*>(base + L::bsum);
    auto* raw_p = reinterpret_cast<float*>(base + L::raw_p);
    auto* P = reinterpret_cast<float4*>(base + L::P);
    auto* PS = reinterpret_cast<float4*>(base + L::PS);
    auto* raw_q = reinterpret_cast<float*>(base + L::raw_q);
    auto* Q = reinterpret_cast<float4*>(base + L::Q);
    auto* cls = reinterpret_cast<int*>(base + L::cls);
    auto* off = reinterpret_cast<int*>(base + L::off);
    auto* tperm = reinterpret_cast<int*>(base + L::tperm);
    auto* tkey = reinterpret_cast<int*>(base + L::tkey);
    auto* perm = reinterpret_cast<int*>(base + L::perm);
    auto* mperm = reinterpret_cast<int*>(base + L::mperm);
    auto* mkey = reinterpret_cast<int*>(base + L::mkey);
    auto* bigb = reinterpret_cast<unsigned char*>(base + L::bigb);
    auto* cnt = reinterpret_cast<int*>(base + L::cnt);
    auto* off64 = reinterpret_cast<long long*>(base + L::off64);
    auto* out = reinterpret_cast<int*>(base + L::out);
    unsigned htcap = 1;
    while ((long long)htcap < 2 * ncell) htcap <<= 1;  // <= 2^19 = L::HT
    hipError_t e = hipSuccess;
    auto scan = [&](int* data, int len) {
      const int nb = cdiv(len, 1024);
      hipError_t r = na_launch(k_scan_reduce, nb, 256, (const int*)data, len, bsum);
      if (r == hipSuccess) r = na_launch(k_scan_bsum, 1, 1024, bsum, nb);
      if (r == hipSuccess) r = na_launch(k_scan_down, nb, 256, data, len, (const int*)bsum);
      return r;
    };
    // inputs to the device, repacked as float4
    if (n > 0) e = hipMemcpy(raw_p, pos3, sizeof(float) * 3 * (size_t)n, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(raw_q, query3, sizeof(float) * 3 * (size_t)nq, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = na_launch(k_pack_float4, cdiv(n, 256), 256, (const float*)raw_p, n, P);
    if (e == hipSuccess) e = na_launch(k_pack_float4, cdiv(nq, 256), 256, (const float*)raw_q, nq, Q);
    // alias-class table of the window (the reference's hash buckets)
    if (e == hipSuccess) e = na_launch(k_ht_clear, cdiv(htcap, 256), 256, htk, htv, (int)htcap);
    if (e == hipSuccess)
      e = na_launch(k_ht_insert, cdiv(nc, 256), 256, htk, htv, htcap - 1, wlo[0], wlo[1], wlo[2], wdim[0], wdim[1],
                    nc);
    if (e == hipSuccess)
      e = na_launch(k_ht_lookup, cdiv(nc, 256), 256, (const unsigned long long*)htk, (const int*)htv, htcap - 1,
                    wlo[0], wlo[1], wlo[2], wdim[0], wdim[1], nc, 0, clsmap);
    // points sorted into classes, ascending index inside a class (bucket order)
    if (e == hipSuccess) e = na_launch(k_fill_i32, cdiv(nc + 1, 256), 256, cst, 0, nc + 1);
    if (e == hipSuccess) e = na_launch(k_aux_bin, cdiv(n, NPX_BIN_BLOCK), NPX_BIN_BLOCK, W, (const float4*)P, n, (const int*)clsmap, cst, cls, off);
    if (e == hipSuccess) e = scan(cst, nc + 1);
    if (e == hipSuccess)
      e = na_launch(k_aux_scatter, cdiv(n, 256), 256, n, (const int*)cst, (const int*)cls, (const int*)off,
                    (const int*)nullptr, tperm, tkey, (int*)nullptr, (int*)nullptr, 31);
    // rank inside classes (off now holds each sorted slot's class), then the
    // merge launches of classes spanning several 256-slot blocks
    const int nblk = cdiv(n, NPX_BLOCK);
    if (e == hipSuccess)
      e = na_launch(k_aux_rank_sort, nblk, NPX_BLOCK, nc, (const int*)cst, (const int*)cls, tperm, tkey, bigb,
                    (const int*)nullptr, 31);
    if (e == hipSuccess)
      e = na_launch(k_aux_rank, nblk, NPX_BLOCK, n, nc, (const int*)cst, (const int*)cls, (const int*)tperm,
                    (const int*)tkey, (const float4*)P, perm, PS, off, (const int*)nullptr, 31,
                    (unsigned char*)nullptr, (int*)nullptr);
    for (int lv = 0; e == hipSuccess && (1 << lv) < nblk + 1; lv += NPX_MERGE_LOG) {
      const bool even = ((lv / NPX_MERGE_LOG) & 1) == 0;
      e = na_launch(k_aux_rank_merge, nblk, NPX_BLOCK, nc, (const int*)cst, (const int*)off,
                    (const unsigned char*)bigb, lv, (const int*)(even ? tperm : mperm), (const int*)(even ? tkey : mkey),
                    even ? mperm : tperm, even ? mkey : tkey, (const float4*)P, perm, PS, (const int*)nullptr, 31);
    }
    // candidate counts per query, then their exclusive prefix (cnt[nq] = total)
    if (e == hipSuccess) e = na_launch(k_fill_i32, cdiv(nq + 1, 256), 256, cnt, 0, nq + 1);
    if (e == hipSuccess)
      e = na_launch(k_query_count, cdiv(nq, 256), 256, W, (const float4*)Q, nq, (const int*)clsmap, (const int*)cst, cnt);
    if (e == hipSuccess && counts) e = hipMemcpy(counts, cnt, sizeof(int) * (size_t)nq, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = scan(cnt, nq + 1);
    int t = 0;
    if (e == hipSuccess) e = hipMemcpy(&t, cnt + nq, sizeof(int), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return bad(NPX_EDEVICE, hipGetErrorString(e));
    *total = t;
    if (!lists || cap < t || t == 0) return NPX_OK;
    if (t > NPX_NA_MAX_LIST) return bad(NPX_ENOMEM, "candidate lists longer than 2^21 entries");
    e = na_launch(k_widen_i32, cdiv(nq, 256), 256, (const int*)cnt, nq, off64);
    if (e == hipSuccess)
      e = na_launch(k_query_lists, cdiv(nq, 256), 256, W, (const float4*)Q, nq, (const int*)clsmap, (const int*)cst,
                    (const int*)perm, (const long long*)off64, out);
    if (e == hipSuccess) e = hipMemcpy(lists, out, sizeof(int) * (size_t)t, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return bad(NPX_EDEVICE, hipGetErrorString(e));
    return NPX_OK;
  });
}

namespace {
// Per-shape loop bounds and constants with the samplers' float expressions
// (boundary.cpp:15-127).
int prepare_shapes(const npx_shape* shapes, int nshapes, float spacing, float extent, std::vector<ShapeS>& out,
                   long long* total) {
  out.assign((size_t)nshapes, ShapeS{});
  long long t = 0;
  for (int k = 0; k < nshapes; ++k) {
    const npx_shape& a = shapes[k];
    ShapeS& S = out[(size_t)k];
    S.type = a.type;
    S.body = a.body;
    S.first = (int)t;
    for (int d = 0; d < 3; ++d) {
      S.pos[d] = a.position[d];
      S.he[d] = a.half_extents[d];
      S.nrm[d] = a.normal[d];
    }
    for (int d = 0; d < 4; ++d) S.rot[d] = a.rotation[d];
    long long cnt = 0;
    if (a.type == NPX_SHAPE_BOX) {
      for (int axis = 0; axis < 3; ++axis) {
        const int a1 = (axis + 1) % 3, a2 = (axis + 2) % 3;
        S.n1[axis] = std::max(1, static_cast<int>(std::floor(2.0f * a.half_extents[a1] / spacing)) + 1);
        S.n2[axis] = std::max(1, static_cast<int>(std::floor(2.0f * a.half_extents[a2] / spacing)) + 1);
        S.axis_first[axis] = (int)cnt;
        cnt += 2ll * S.n1[axis] * S.n2[axis];
      }
    } else if (a.type == NPX_SHAPE_SPHERE) {
      const float r = a.radius;
      const float pi = 3.14159265358979323846f;
      const float area = 4.0f * pi * r * r;
      S.n = std::max(12, static_cast<int>(area / (spacing * spacing)));
      const float golden_ratio = (1.0f + std::sqrt(5.0f)) / 2.0f;
      S.golden_angle = 2.0f * pi / (golden_ratio * golden_ratio);
      S.r = r;
      cnt = S.n;
    } else if (a.type == NPX_SHAPE_PLANE) {
      const float* n = a.normal;
      float t1[3];
      if (std::fabs(n[1]) < 0.9f) {  // (0, 1, 0) x n
        t1[0] = 1.0f * n[2] - 0.0f * n[1];
        t1[1] = 0.0f * n[0] - 0.0f * n[2];
        t1[2] = 0.0f * n[1] - 1.0f * n[0];
      } else {  // (1, 0, 0) x n
        t1[0] = 0.0f * n[2] - 0.0f * n[1];
        t1[1] = 0.0f * n[0] - 1.0f * n[2];
        t1[2] = 1.0f * n[1] - 0.0f * n[0];
      }
      auto normalize = [](float* v) {
        const float z = v[0] * v[0] + (v[1] * v[1] + v[2] * v[2]);
        if (z > 0.0f) {
          const float l = std::sqrt(z);
          for (int d = 0; d < 3; ++d) v[d] = v[d] / l;
        }
      };
      normalize(t1);
      float t2[3] = {n[1] * t1[2] - n[2] * t1[1], n[2] * t1[0] - n[0] * t1[2], n[0] * t1[1] - n[1] * t1[0]};
      normalize(t2);
      for (int d = 0; d < 3; ++d) {
        S.t1[d] = t1[d];
        S.t2[d] = t2[d];
      }
      S.offset = a.offset;
      S.extent = extent;
      S.n = static_cast<int>(std::floor(2.0f * extent / spacing)) + 1;
      cnt = (long long)S.n * S.n;
    } else {
      return NPX_EINVAL;
    }
    S.count = (int)cnt;
    t += cnt;
  }
  *total = t;
  return NPX_OK;
}
}  // namespace

int npx_sample_boundaries(const npx_shape* shapes, int32_t nshapes, float spacing, float plane_extent,
                          float* local3, int32_t* body, int64_t cap, int64_t* count) {
  return guard(nullptr, __func__, [&]() -> int {
    if (!count || nshapes < 0 || (nshapes > 0 && !shapes) || !(spacing > 0.0f)) {
      g_create_error = "npx_sample_boundaries: bad arguments";
      return NPX_EINVAL;
    }
    std::vector<ShapeS> hs;
    long long total = 0;
    if (prepare_shapes(shapes, nshapes, spacing, plane_extent, hs, &total) != NPX_OK) {
      g_create_error = "npx_sample_boundaries: unknown shape type";
      return NPX_EINVAL;
    }
    if (total > (1ll << 30)) {
      g_create_error = "npx_sample_boundaries: too many boundary particles";
      return NPX_EINVAL;
    }
    *count = total;
    if (total == 0 || total > cap) return NPX_OK;
    if (!local3 || !body) {
      g_create_error = "npx_sample_boundaries: output is NULL";
      return NPX_EINVAL;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
      g_create_error = "no HIP device available";
      return NPX_EDEVICE;
    }
    // The sphere samples' trigonometry, with the reference's float calls
    // (sample_sphere_boundary, boundary.cpp:55-91): the theta argument
    // 1 - 2 (i + 0.5) / n and phi = golden_angle * i in float, then acos,
    // sin and cos of float.
    std::vector<float4> trig((size_t)total, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
    for (const ShapeS& S : hs) {
      if (S.type != NPX_SHAPE_SPHERE) continue;
      for (int i = 0; i < S.n; ++i) {
        const float theta = std::acos(1.0f - 2.0f * ((float)i + 0.5f) / (float)S.n);
        const float phi = S.golden_angle * (float)i;
        trig[(size_t)S.first + i] = make_float4(std::sin(theta), std::cos(theta), std::sin(phi), std::cos(phi));
      }
    }
    ShapeS* ds = nullptr;
    float4* dt = nullptr;
    float* dl = nullptr;
    int* db = nullptr;
    hipError_t e = hipMalloc((void**)&ds, sizeof(ShapeS) * hs.size());
    if (e == hipSuccess) e = hipMalloc((void**)&dt, sizeof(float4) * (size_t)total);
    if (e == hipSuccess) e = hipMalloc((void**)&dl, sizeof(float) * 3 * (size_t)total);
    if (e == hipSuccess) e = hipMalloc((void**)&db, sizeof(int) * (size_t)total);
    if (e == hipSuccess) e = hipMemcpy(ds, hs.data(), sizeof(ShapeS) * hs.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dt, trig.data(), sizeof(float4) * (size_t)total, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
      const int n = (int)total;
      hipLaunchKernelGGL(k_sample_boundaries, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, 0, (const ShapeS*)ds,
                         nshapes, n, spacing, (const float4*)dt, dl, db);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(local3, dl, sizeof(float) * 3 * (size_t)total, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(body, db, sizeof(int) * (size_t)total, hipMemcpyDeviceToHost);
    if (ds) (void)hipFree(ds);
    if (dt) (void)hipFree(dt);
    if (dl) (void)hipFree(dl);
    if (db) (void)hipFree(db);
    if (e != hipSuccess) {
      g_create_error = std::string("npx_sample_boundaries: ") + hipGetErrorString(e);
      return NPX_EDEVICE;
    }
    return NPX_OK;
  });
}

int npx_set_boundary(npx_ctx* ctx, const float* local3, const int32_t* body, const float* psi, int32_t nb,
                     int32_t nbodies) {
  return guard(ctx, __func__, [&]() -> int {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c) return NPX_EINVAL;
    if (c->batch) return c->fail(NPX_ESTATE, "boundary coupling is for single-scene contexts");
    return set_boundary(c, local3, body, psi, nb, nbodies);
  });
}

int npx_compute_boundary_volumes(npx_ctx* ctx, const float* pose7, int32_t nbodies, float* psi_out) {
  return guard(ctx, __func__, [&]() -> int {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c) return NPX_EINVAL;
    return compute_volumes(c, pose7, nbodies, psi_out);
  });
}

int npx_boundary_world_positions(npx_ctx* ctx, const float* pose7, int32_t nbodies, float* out3) {
  return guard(ctx, __func__, [&]() -> int {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c) return NPX_EINVAL;
    if (c->nb == 0) return NPX_OK;
    NPX_TRY(bpos(c, pose7, nbodies));
    if (out3) {
      std::vector<float4> bp(c->nb);
      NPX_HIP_C(c, hipMemcpyAsync(bp.data(), c->BP.p, sizeof(float4) * c->nb, hipMemcpyDeviceToHost, c->stream));
      NPX_HIP_C(c, hipStreamSynchronize(c->stream));
      for (int b = 0; b < c->nb; ++b) {
        out3[3 * b] = bp[b].x;
        out3[3 * b + 1] = bp[b].y;
        out3[3 * b + 2] = bp[b].z;
      }
    }
    return NPX_OK;
  });
}

int npx_apply_boundary_density(npx_ctx* ctx) {
  return guard(ctx, __func__, [&]() -> int {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c) return NPX_EINVAL;
    return boundary_density(c);
  });
}

int npx_coupling_forces(npx_ctx* ctx, const float* pose7, const uint8_t* is_static, int32_t nbodies, float* force3,
                        float* torque3) {
  return guard(ctx, __func__, [&]() -> int {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c) return NPX_EINVAL;
    return coupling(c, pose7, is_static, nbodies, force3, torque3);
  });
}

// ------------------------------------------------------------- batch
int npx_batch_create(int32_t slot_capacity, int32_t cell_capacity, int32_t max_scenes, uint32_t flags,
                     npx_ctx** out) {
  return guard(nullptr, __func__, [&]() -> int {
    if (!out || slot_capacity <= 0 || cell_capacity <= 0 || max_scenes <= 0) {
      g_create_error = "npx_batch_create: capacities must be positive";
      return NPX_EINVAL;
    }
    if (slot_capacity > NPX_MAX_SLOTS || cell_capacity > (1 << 30) || max_scenes > (1 << 20)) {
      g_create_error = "npx_batch_create: capacity beyond the limits (2^28 slots, 2^30 cells, 2^20 scenes)";
      return NPX_EINVAL;
    }
    *out = nullptr;
    Ctx* c = new (std::nothrow) Ctx();
    if (!c) {
      g_create_error = "host memory allocation failed";
      return NPX_ENOMEM;
    }
    int r = c->init(slot_capacity, cell_capacity, max_scenes, flags, true);
    if (r != NPX_OK) {
      g_create_error = c->err;
      destroy(c);
      return r;
    }
    *out = reinterpret_cast<npx_ctx*>(c);
    return NPX_OK;
  });
}

int npx_scene_add(npx_ctx* ctx, const npx_pbf_params* params, float dt, const float gravity[3], float particle_mass,
                  const float* pos3, const float* vel3, int32_t n, int32_t* scene_id) {
  return guard(ctx, __func__, [&]() -> int {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c) return NPX_EINVAL;
    if (!c->batch) return c->fail(NPX_ESTATE, "npx_scene_add needs a batch context");
    NPX_TRY(check_params(c, params));
    if (!params->use_domain_bounds) return c->fail(NPX_EINVAL, "batched scenes need use_domain_bounds");
    if (n < 0 || (n > 0 && !pos3) || !gravity) return c->fail(NPX_EINVAL, "bad scene arrays");
    int sid = -1;
    for (int s = 0; s < c->max_scenes; ++s)
      if (!c->scenes[s].used) {
        sid = s;
        break;
      }
    if (sid < 0) return c->fail(NPX_ENOMEM, "all %d scene slots are in use", c->max_scenes);
    SceneH H;
    H.prm = *params;
    H.dt = dt;
    H.mass = particle_mass;
    std::memcpy(H.g, gravity, sizeof(float) * 3);
    H.n = n;
    H.seg_blocks = cdiv(std::max(n, 1), NPX_BLOCK);
    const int b0 = c->slot_alloc.alloc(H.seg_blocks);
    if (b0 < 0) return c->fail(NPX_ENOMEM, "slot capacity exhausted (%d particles requested)", n);
    H.seg_off = b0 * NPX_BLOCK;
    int r = c->scene_window(H);
    if (r != NPX_OK) {
      c->slot_alloc.release(b0, H.seg_blocks);
      return r;
    }
    H.used = true;
    c->scenes[sid] = H;
    for (int b = 0; b < H.seg_blocks; ++b) c->h_blk_scene[b0 + b] = (b < cdiv(n, NPX_BLOCK)) ? sid : -1;
    c->table_dirty = true;
    NPX_TRY(c->upload_particles(c->scenes[sid], pos3, vel3));
    NPX_TRY(c->upload_tables());
    NPX_TRY(c->reset_list_stats(sid));
    if (scene_id) *scene_id = sid;
    return NPX_OK;
  });
}

int npx_scene_remove(npx_ctx* ctx, int32_t scene_id) {
  return guard(ctx, __func__, [&]() -> int {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c) return NPX_EINVAL;
    if (scene_id < 0 || scene_id >= c->max_scenes || !c->scenes[scene_id].used)
      return c->fail(NPX_EINVAL, "no scene %d", scene_id);
    SceneH& H = c->scenes[scene_id];
    NPX_HIP_C(c, hipStreamSynchronize(c->stream));
    const int b0 = H.seg_off / NPX_BLOCK;
    for (int b = 0; b < H.seg_blocks; ++b) c->h_blk_scene[b0 + b] = -1;
    c->slot_alloc.release(b0, H.seg_blocks);
    if (H.window_valid) c->cls_alloc.release(H.cls_base, H.ncell);
    NPX_TRY(c->release_boundary(scene_id));
    H = SceneH();
    c->table_dirty = true;
    return c->upload_tables();
  });
}

int npx_scene_get(npx_ctx* ctx, int32_t scene_id, float* pos3, float* vel3, float* rho, float* lambda) {
  return guard(ctx, __func__, [&]() -> int {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c) return NPX_EINVAL;
    if (scene_id < 0 || scene_id >= c->max_scenes || !c->scenes[scene_id].used)
      return c->fail(NPX_EINVAL, "no scene %d", scene_id);
    return c->download(c->scenes[scene_id], pos3, vel3, rho, lambda);
  });
}

int npx_scene_set_boundary(npx_ctx* ctx, int32_t scene_id, const float* local3, const int32_t* body,
                           const float* psi, int32_t nb, int32_t nbodies, const uint8_t* is_static) {
  return guard(ctx, __func__, [&]() -> int {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c) return NPX_EINVAL;
    if (!c->batch) return c->fail(NPX_ESTATE, "npx_scene_set_boundary needs a batch context");
    if (scene_id < 0 || scene_id >= c->max_scenes || !c->scenes[scene_id].used)
      return c->fail(NPX_EINVAL, "no scene %d", scene_id);
    if (nb < 0 || nbodies < 0) return c->fail(NPX_EINVAL, "negative boundary size");
    if (nb > 0 && (!local3 || !body || !psi)) return c->fail(NPX_EINVAL, "boundary arrays are NULL");
    if (nbodies > 0 && !is_static) return c->fail(NPX_EINVAL, "is_static is NULL");
    NPX_HIP_C(c, hipStreamSynchronize(c->stream));
    if (c->sbnd.size() < (size_t)c->max_scenes) c->sbnd.resize(c->max_scenes);
    NPX_TRY(c->release_boundary(scene_id));
    Ctx::SceneBnd& B = c->sbnd[scene_id];
    B.local3.assign(local3, local3 + 3 * (size_t)nb);
    B.body.assign(body, body + nb);
    B.psi.assign(psi, psi + nb);
    B.is_static.assign(is_static, is_static + nbodies);
    B.nbodies = nbodies;
    c->scenes[scene_id].nbnd = nb;
    c->body_dirty = true;
    NPX_TRY(c->place_boundary(scene_id));
    c->table_dirty = true;
    return c->upload_tables();
  });
}

int32_t npx_batch_num_bodies(npx_ctx* ctx) {
  return guard(ctx, __func__, [&]() -> int {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c) return -1;
    if (c->body_dirty && c->rebuild_bodies() != NPX_OK) return -1;
    return c->bb_bodies;
  });
}

int npx_batch_coupling(npx_ctx* ctx, const float* pose7, int32_t nbodies, float* out6) {
  return guard(ctx, __func__, [&]() -> int {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c) return NPX_EINVAL;
    if (!c->batch) return c->fail(NPX_ESTATE, "npx_batch_coupling needs a batch context");
    return c->batch_coupling(pose7, nbodies, out6);
  });
}

int npx_batch_step(npx_ctx* ctx, int32_t nsteps) {
  return guard(ctx, __func__, [&]() -> int {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c) return NPX_EINVAL;
    return c->step_n(nsteps);
  });
}

int npx_set_tuning(npx_ctx* ctx, int32_t block_size, int32_t xcd_remap, float skin_fraction) {
  return guard(ctx, __func__, [&]() -> int {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c) return NPX_EINVAL;
    if (block_size != 64 && block_size != 128 && block_size != 256)
      return c->fail(NPX_EINVAL, "block_size must be 64, 128 or 256");
    if (xcd_remap < 0 || xcd_remap > 4096) return c->fail(NPX_EINVAL, "xcd_remap must be in [0, 4096]");
    if (!(skin_fraction >= 0.02f && skin_fraction <= 0.5f))
      return c->fail(NPX_EINVAL, "skin_fraction must be in [0.02, 0.5]");
    NPX_HIP_C(c, hipStreamSynchronize(c->stream));
    c->nbr_block = block_size;
    c->xcd_remap = xcd_remap;
    c->skin_frac = skin_fraction;
    c->table_dirty = true;
    return NPX_OK;
  });
}

int npx_set_persistent(npx_ctx* ctx, int32_t blocks) {
  return guard(ctx, __func__, [&]() -> int {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c) return NPX_EINVAL;
    if (blocks != 0 && (blocks < 8 || blocks % 8 != 0 || blocks > (1 << 22)))
      return c->fail(NPX_EINVAL, "persistent block count must be 0 or a multiple of 8 in [8, 2^22]");
    NPX_HIP_C(c, hipStreamSynchronize(c->stream));
    c->persist_blocks = blocks;
    c->table_dirty = true;
    return NPX_OK;
  });
}

int npx_set_list_policy(npx_ctx* ctx, int32_t policy) {
  return guard(ctx, __func__, [&]() -> int {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c) return NPX_EINVAL;
    if (policy != 0 && policy != 1) return c->fail(NPX_EINVAL, "list policy must be 0 or 1");
    NPX_HIP_C(c, hipStreamSynchronize(c->stream));
    c->list_policy = policy;
    // no step has counted list statistics under the new policy yet
    NPX_HIP_C(c, hipMemsetAsync(c->LST.p, 0, sizeof(int) * NPX_LST_STRIDE * (size_t)c->max_scenes, c->stream));
    NPX_HIP_C(c, hipStreamSynchronize(c->stream));
    return NPX_OK;
  });
}

int npx_set_rank_split(npx_ctx* ctx, int32_t huge_levels) {
  return guard(ctx, __func__, [&]() -> int {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c) return NPX_EINVAL;
    if (huge_levels < 0 || huge_levels > 24) return c->fail(NPX_EINVAL, "huge_levels must be in [0, 24]");
    NPX_HIP_C(c, hipStreamSynchronize(c->stream));
    c->huge_levels = huge_levels;
    return NPX_OK;
  });
}

int npx_batch_step_timed(npx_ctx* ctx, int32_t nsteps, float* step_ms) {
  return guard(ctx, __func__, [&]() -> int {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c || nsteps < 0) return NPX_EINVAL;
    return c->step_timed(nsteps, step_ms);
  });
}

int npx_scene_status(npx_ctx* ctx, int32_t* status, int32_t max_scenes) {
  return guard(ctx, __func__, [&]() -> int {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c || !status) return NPX_EINVAL;
    return c->scene_status(status, max_scenes);
  });
}

int64_t npx_scene_bytes(const npx_pbf_params* params, int32_t n) {
  if (!params || !(params->kernel_radius > 0.0f)) return -1;
  int wlo[3], wdim[3];
  Ctx::window_from_bounds(1.0f / params->kernel_radius, params->domain_lower, params->domain_upper,
                          NPX_WINDOW_MARGIN, wlo, wdim);
  const long long cells = (long long)wdim[0] * wdim[1] * wdim[2];
  const long long slots = (long long)cdiv(std::max(n, 1), NPX_BLOCK) * NPX_BLOCK;
  // 9 float4 + 5 float + 17 int + 2 byte arrays, a huge-class count row
  // entry per slot and the skin and in-range lists per slot; class map,
  // start, displacement maximum and run flag per cell
  return slots * (9 * 16 + 5 * 4 + 17 * 4 + 2 + NPX_HUGE_MAX * 4 / NPX_BLOCK + 8 * NPX_SKIN_CAP) + cells * 16;
}

int npx_set_kernel_timing(npx_ctx* ctx, int32_t enable) {
  return guard(ctx, __func__, [&]() -> int {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c) return NPX_EINVAL;
    NPX_HIP_C(c, hipStreamSynchronize(c->stream));
    for (auto& t : c->timers) {
      c->ev_pool.push_back(t.a);
      c->ev_pool.push_back(t.b);
    }
    c->timers.clear();
    c->timing = enable != 0;
    return NPX_OK;
  });
}

int npx_kernel_times(npx_ctx* ctx, char* names, int32_t names_cap, float* ms, int32_t* launches,
                     int32_t max_kernels, int32_t* count) {
  return guard(ctx, __func__, [&]() -> int {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c || !count) return NPX_EINVAL;
    NPX_HIP_C(c, hipStreamSynchronize(c->stream));
    std::vector<std::string> order;
    std::map<std::string, std::pair<double, int>> acc;
    for (auto& t : c->timers) {
      float e = 0.0f;
      NPX_HIP_C(c, hipEventElapsedTime(&e, t.a, t.b));
      if (!acc.count(t.name)) order.push_back(t.name);
      acc[t.name].first += e;
      acc[t.name].second += 1;
    }
    *count = (int32_t)order.size();
    std::string joined;
    for (size_t k = 0; k < order.size(); ++k) {
      if ((int)k < max_kernels) {
        if (ms) ms[k] = (float)acc[order[k]].first;
        if (launches) launches[k] = acc[order[k]].second;
      }
      joined += order[k];
      joined += '\n';
    }
    if (names && names_cap > 0) {
      std::strncpy(names, joined.c_str(), names_cap - 1);
      names[names_cap - 1] = '\0';
    }
    return NPX_OK;
  });
}

int npx_kernel_trace(npx_ctx* ctx, char* names, int32_t names_cap, float* start_ms, float* dur_ms,
                     int32_t max_launches, int32_t* count) {
  return guard(ctx, __func__, [&]() -> int {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c || !count) return NPX_EINVAL;
    NPX_HIP_C(c, hipStreamSynchronize(c->stream));
    std::string joined;
    int k = 0;
    for (auto& t : c->timers) {
      if (k < max_launches) {
        float st = 0.0f, d = 0.0f;
        NPX_HIP_C(c, hipEventElapsedTime(&st, c->timers.front().a, t.a));
        NPX_HIP_C(c, hipEventElapsedTime(&d, t.a, t.b));
        if (start_ms) start_ms[k] = st;
        if (dur_ms) dur_ms[k] = d;
        joined += t.name;
        joined += '\n';
      }
      ++k;
    }
    *count = k;
    if (names && names_cap > 0) {
      std::strncpy(names, joined.c_str(), names_cap - 1);
      names[names_cap - 1] = '\0';
    }
    return NPX_OK;
  });
}

}  // extern "C"
