// Obstacle skinning and cloth-to-body bindings: set-up, scheduling and the C entry points.
//
// At xrt_start (ClothSolverBase::SetupBinding and the pipelines' set-up,
// xrtailor/src/pipeline/impl/smpl/ClothSolver.cpp:94-107, ClothSolverBase.cpp:448-451):
// the obstacle tree is built from its rest state with the bvh tolerance (UpdateData +
// Construct), every cloth vertex of an instance with bindings gets its nearest obstacle
// face and barycentric weights (QueryNearestBarycentricClothToObstacle), then the bindings
// are generated.  Every iteration ends with BindingConstraint::Solve against the obstacle's
// frame positions (ClothSolverBase.cpp:237).  LBS (body.cu:39-58) writes the obstacle's
// frame positions from joint transforms, like xrt_update_obstacle.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "binding.h"
#include "context.h"

namespace xrt {

namespace {
constexpr int kBlock = 256;

int Blocks(int n) { return std::max(0, (n + kBlock - 1) / kBlock); }

DevMap Map(const DevBuf<int32_t>& maps, const MapRef& r) {
  DevMap m;
  m.off = maps.ptr + r.off;
  m.n = r.n;
  m.blocks = r.blocks;
  return m;
}

int CheckInstance(xrt_ctx* c, int32_t inst, bool started) {
  XRT_REQUIRE(c != nullptr, XRT_ERR_INVALID_ARG, "null context");
  XRT_REQUIRE(inst >= 0 && inst < static_cast<int32_t>(c->inst.size()), XRT_ERR_INVALID_ARG,
              "instance index out of range");
  XRT_REQUIRE(c->started == started, XRT_ERR_STATE,
              started ? "xrt_start has not been called" : "call before xrt_start");
  return XRT_OK;
}

// the obstacle tree as the set-up query sees it: rest positions, bvh tolerance along the
// vertex normals, no CCD; then the nearest record of every cloth vertex of the instance
int NearestRecords(xrt_ctx* ctx, int inst) {
  ContactData& c = *ctx->contact[inst];
  const InstDesc& d = ctx->inst[inst].desc;
  hipStream_t s = ctx->stream;
  BindingData& B = ctx->binding;
  XRT_CHECK_HIP(VertexNormals(c.o_x0.ptr, c.o_tris.ptr, c.o_vptr.ptr, c.o_corner.ptr, c.obs_nv,
                              c.o_n1.ptr, s));
  XRT_CHECK_HIP(LbvhBuild(c.obs_bvh, c.o_x0.ptr, nullptr, c.o_n1.ptr, ctx->params.bvh_tolerance, 0,
                          0.0f, s));
  XRT_CHECK_HIP(hipMemsetAsync(B.flag.ptr, 0, sizeof(int32_t), s));
  XRT_CHECK_HIP(LaunchNearestBary(s, c.obs_bvh.view(), ctx->x0.ptr + d.vbeg, d.nv,
                                  B.near_idx.ptr + d.vbeg, B.near_uvw.ptr + d.vbeg, B.flag.ptr));
  int32_t flag = 0;
  XRT_CHECK_HIP(hipMemcpyAsync(&flag, B.flag.ptr, sizeof(flag), hipMemcpyDeviceToHost, s));
  XRT_CHECK_HIP(hipStreamSynchronize(s));
  XRT_REQUIRE(flag == 0, XRT_ERR_CAPACITY, "nearest-face search exceeded its 64-entry stack");
  return XRT_OK;
}

// The transforms go through the pinned staging slots, so the call returns without waiting
// for the stream: the host can queue the next frame while the current one runs.
int Skin(xrt_ctx* ctx, const MapRef& r, const float* transforms, size_t count, int inst) {
  BindingData& B = ctx->binding;
  hipStream_t s = ctx->stream;
  for (size_t k = 0; k < count; ++k)
    XRT_REQUIRE(std::isfinite(transforms[k]), XRT_ERR_NOT_FINITE, "non-finite joint transform");
  const size_t off = inst < 0 ? 0 : static_cast<size_t>(B.joint_off[inst]) * 12;
  SkinStage& st = B.stage;
  const size_t per = 12 * static_cast<size_t>(B.total_joints);
  if (st.per < per) {
    st.Release();
    for (int k = 0; k < 2; ++k)
      XRT_CHECK_HIP(hipEventCreateWithFlags(&st.ev[k], hipEventDisableTiming));
    XRT_CHECK_HIP(hipHostMalloc(reinterpret_cast<void**>(&st.host), sizeof(float) * 2 * per,
                                hipHostMallocDefault));
    st.per = per;
  }
  const int slot = st.slot;
  if (!st.fresh[slot]) XRT_CHECK_HIP(hipEventSynchronize(st.ev[slot]));
  float* h = st.host + static_cast<size_t>(slot) * st.per + off;
  std::memcpy(h, transforms, sizeof(float) * count);
  XRT_CHECK_HIP(hipMemcpyAsync(B.transforms.ptr + off, h, sizeof(float) * count,
                               hipMemcpyHostToDevice, s));
  XRT_CHECK_HIP(hipEventRecord(st.ev[slot], s));
  st.fresh[slot] = false;
  st.slot = slot ^ 1;
  XRT_CHECK_HIP(LaunchLbs(s, B.skin_dev.ptr, B.transforms.ptr, Map(B.skin_maps, r)));
  for (size_t k = 0; k < ctx->inst.size(); ++k)
    if (B.joint_off[k] >= 0 && (inst < 0 || static_cast<int>(k) == inst))
      ctx->contact[k]->obstacle_moving = true;
  return XRT_OK;
}
}  // namespace

void SkinStage::Release() {
  for (int k = 0; k < 2; ++k) {
    if (ev[k]) {
      (void)hipEventSynchronize(ev[k]);
      (void)hipEventDestroy(ev[k]);
    }
    ev[k] = nullptr;
    fresh[k] = true;
  }
  if (host) (void)hipHostFree(host);
  host = nullptr;
  per = 0;
  slot = 0;
}

bool BindingsOn(const xrt_ctx* ctx) { return ctx->binding.total > 0; }

int BindingStart(xrt_ctx* ctx) {
  BindingData& B = ctx->binding;
  hipStream_t s = ctx->stream;
  const int n = static_cast<int>(ctx->inst.size());
  B.inst.assign(n, BindInstDev{});
  B.joint_off.assign(n, -1);
  B.skin.clear();
  B.total = 0;
  B.total_joints = 0;
  std::vector<int> bblocks(n, 0), sblocks(n, 0);
  int skinned = 0;
  for (int k = 0; k < n; ++k) {
    const HostInstance& h = ctx->inst[k];
    BindInstDev& bi = B.inst[k];
    bi.bbeg = B.total;
    bi.nb = static_cast<int>(h.bind_vert.size());
    bi.vbeg = h.desc.vbeg;
    bi.o_x = ctx->contact[k] ? ctx->contact[k]->o_x.ptr : nullptr;
    B.total += bi.nb;
    bblocks[k] = Blocks(bi.nb);
    if (h.n_joints > 0) {
      B.joint_off[k] = B.total_joints;
      B.total_joints += h.n_joints;
      sblocks[k] = Blocks(h.obs_nv());
      ++skinned;
    }
  }
  if (B.total == 0 && skinned == 0) return XRT_OK;
  {  // maps
    std::vector<int32_t> maps, smaps;
    B.all = AppendBlockMap(maps, 0, n, bblocks);
    B.skin_all = AppendBlockMap(smaps, 0, n, sblocks);
    B.one.resize(n);
    B.skin_one.resize(n);
    for (int k = 0; k < n; ++k) {
      B.one[k] = AppendBlockMap(maps, k, k + 1, bblocks);
      B.skin_one[k] = AppendBlockMap(smaps, k, k + 1, sblocks);
    }
    XRT_CHECK_HIP(B.maps.Upload(maps, s));
    XRT_CHECK_HIP(B.skin_maps.Upload(smaps, s));
  }
  if (B.total > 0) {
    const size_t N = static_cast<size_t>(ctx->wide.n_total);
    XRT_CHECK_HIP(B.dev.Upload(B.inst, s));
    XRT_CHECK_HIP(B.items.Alloc(B.total));
    XRT_CHECK_HIP(B.near_idx.Alloc(N));
    XRT_CHECK_HIP(B.near_uvw.Alloc(N));
    XRT_CHECK_HIP(B.flag.Alloc(1));
    for (int k = 0; k < n; ++k) {
      const HostInstance& h = ctx->inst[k];
      if (h.bind_vert.empty()) continue;
      int rc = NearestRecords(ctx, k);
      if (rc != XRT_OK) return rc;
      DevBuf<int32_t> vert;
      DevBuf<float> stiff, dist;
      XRT_CHECK_HIP(vert.Upload(h.bind_vert, s));
      XRT_CHECK_HIP(stiff.Upload(h.bind_stiff, s));
      XRT_CHECK_HIP(dist.Upload(h.bind_dist, s));
      XRT_CHECK_HIP(LaunchBindingGenerate(s, vert.ptr, stiff.ptr, dist.ptr, B.inst[k].nb,
                                          ctx->x0.ptr + h.desc.vbeg, B.near_idx.ptr + h.desc.vbeg,
                                          B.near_uvw.ptr + h.desc.vbeg, ctx->contact[k]->o_x0.ptr,
                                          B.items.ptr + B.inst[k].bbeg));
      XRT_CHECK_HIP(hipStreamSynchronize(s));
    }
  }
  if (skinned > 0) {
    std::vector<int32_t> ptr, joint;
    std::vector<float> weight;
    std::vector<float4> rest;
    std::vector<size_t> pbeg(n), jbeg(n), rbeg(n);
    for (int k = 0; k < n; ++k) {
      const HostInstance& h = ctx->inst[k];
      if (h.n_joints <= 0) continue;
      pbeg[k] = ptr.size();
      jbeg[k] = joint.size();
      rbeg[k] = rest.size();
      ptr.insert(ptr.end(), h.skin_ptr.begin(), h.skin_ptr.end());
      joint.insert(joint.end(), h.skin_joint.begin(), h.skin_joint.end());
      weight.insert(weight.end(), h.skin_weight.begin(), h.skin_weight.end());
      for (int v = 0; v < h.obs_nv(); ++v)
        rest.push_back(make_float4(h.obs_pos[3 * v], h.obs_pos[3 * v + 1], h.obs_pos[3 * v + 2], 0.0f));
    }
    XRT_CHECK_HIP(B.sk_ptr.Upload(ptr, s));
    XRT_CHECK_HIP(B.sk_joint.Upload(joint, s));
    XRT_CHECK_HIP(B.sk_weight.Upload(weight, s));
    XRT_CHECK_HIP(B.sk_rest.Upload(rest, s));
    {  // identity transforms until the first skinning call (the rest pose)
      std::vector<float> eye(12 * static_cast<size_t>(B.total_joints), 0.0f);
      for (int j = 0; j < B.total_joints; ++j) eye[12 * j] = eye[12 * j + 5] = eye[12 * j + 10] = 1.0f;
      XRT_CHECK_HIP(B.transforms.Upload(eye, s));
    }
    B.skin.assign(n, SkinInstDev{});
    for (int k = 0; k < n; ++k) {
      const HostInstance& h = ctx->inst[k];
      if (h.n_joints <= 0) continue;
      SkinInstDev& d = B.skin[k];
      d.nv = h.obs_nv();
      d.nj = h.n_joints;
      d.toff = 12 * B.joint_off[k];
      d.ptr = B.sk_ptr.ptr + pbeg[k];
      d.joint = B.sk_joint.ptr + jbeg[k];
      d.weight = B.sk_weight.ptr + jbeg[k];
      d.rest = B.sk_rest.ptr + rbeg[k];
      d.out = ctx->contact[k]->o_x.ptr;
    }
    XRT_CHECK_HIP(B.skin_dev.Upload(B.skin, s));
  }
  XRT_CHECK_HIP(hipStreamSynchronize(s));
  return XRT_OK;
}

int BindingIter(xrt_ctx* ctx, int inst) {
  BindingData& B = ctx->binding;
  if (B.total == 0) return XRT_OK;
  const MapRef& r = inst < 0 ? B.all : B.one[inst];
  XRT_CHECK_HIP(LaunchBindingSolve(ctx->stream, B.dev.ptr, B.items.ptr, Map(B.maps, r), ctx->x.ptr));
  return XRT_OK;
}

}  // namespace xrt

using namespace xrt;

extern "C" {

int xrt_add_bindings(xrt_ctx* ctx, int32_t inst, const int32_t* idx, const float* stiffness,
                     const float* distance, int32_t n) {
  int rc = CheckInstance(ctx, inst, false);
  if (rc != XRT_OK) return rc;
  XRT_REQUIRE(idx && stiffness && distance && n > 0, XRT_ERR_INVALID_ARG, "empty or null bindings");
  HostInstance& h = ctx->inst[inst];
  XRT_REQUIRE(h.obs_nt() >= 2, XRT_ERR_STATE, "bindings need an obstacle of >= 2 faces (xrt_set_obstacle)");
  std::vector<char> seen(h.nv, 0);
  for (int v : h.bind_vert) seen[v] = 1;
  for (int i = 0; i < n; ++i) {
    XRT_REQUIRE(idx[i] >= 0 && idx[i] < h.nv, XRT_ERR_INVALID_ARG, "binding vertex out of range");
    XRT_REQUIRE(!seen[idx[i]], XRT_ERR_INVALID_ARG, "a vertex is bound twice");
    XRT_REQUIRE(std::isfinite(stiffness[i]) && std::isfinite(distance[i]), XRT_ERR_NOT_FINITE,
                "non-finite binding parameter");
    seen[idx[i]] = 1;
  }
  h.bind_vert.insert(h.bind_vert.end(), idx, idx + n);
  h.bind_stiff.insert(h.bind_stiff.end(), stiffness, stiffness + n);
  h.bind_dist.insert(h.bind_dist.end(), distance, distance + n);
  return XRT_OK;
}

int xrt_set_obstacle_skinning(xrt_ctx* ctx, int32_t inst, const int32_t* ptr, const int32_t* joint,
                              const float* weight, int32_t n_joints) {
  int rc = CheckInstance(ctx, inst, false);
  if (rc != XRT_OK) return rc;
  HostInstance& h = ctx->inst[inst];
  XRT_REQUIRE(h.obs_nt() > 0, XRT_ERR_STATE, "set the obstacle first (xrt_set_obstacle)");
  XRT_REQUIRE(ptr && joint && weight && n_joints > 0 && n_joints <= kMaxJoints, XRT_ERR_INVALID_ARG,
              "null skinning arrays or n_joints outside [1, 256]");
  const int nv = h.obs_nv();
  XRT_REQUIRE(ptr[0] == 0, XRT_ERR_INVALID_ARG, "ptr[0] must be 0");
  for (int v = 0; v < nv; ++v)
    XRT_REQUIRE(ptr[v + 1] >= ptr[v], XRT_ERR_INVALID_ARG, "ptr must be non-decreasing");
  for (int k = 0; k < ptr[nv]; ++k) {
    XRT_REQUIRE(joint[k] >= 0 && joint[k] < n_joints, XRT_ERR_INVALID_ARG, "joint out of range");
    XRT_REQUIRE(std::isfinite(weight[k]), XRT_ERR_NOT_FINITE, "non-finite skinning weight");
  }
  h.skin_ptr.assign(ptr, ptr + nv + 1);
  h.skin_joint.assign(joint, joint + ptr[nv]);
  h.skin_weight.assign(weight, weight + ptr[nv]);
  h.n_joints = n_joints;
  return XRT_OK;
}

int xrt_skin_obstacles(xrt_ctx* ctx, const float* transforms) {
  XRT_REQUIRE(ctx && ctx->started, XRT_ERR_STATE, "xrt_start has not been called");
  XRT_REQUIRE(transforms, XRT_ERR_INVALID_ARG, "null transforms");
  XRT_REQUIRE(ctx->binding.total_joints > 0, XRT_ERR_STATE, "no obstacle has skinning weights");
  return Skin(ctx, ctx->binding.skin_all, transforms, 12 * static_cast<size_t>(ctx->binding.total_joints), -1);
}

int xrt_skin_obstacle(xrt_ctx* ctx, int32_t inst, const float* transforms) {
  int rc = CheckInstance(ctx, inst, true);
  if (rc != XRT_OK) return rc;
  XRT_REQUIRE(transforms, XRT_ERR_INVALID_ARG, "null transforms");
  XRT_REQUIRE(ctx->binding.joint_off[inst] >= 0, XRT_ERR_STATE, "the obstacle has no skinning weights");
  return Skin(ctx, ctx->binding.skin_one[inst], transforms,
              12 * static_cast<size_t>(ctx->inst[inst].n_joints), inst);
}

int xrt_query_nearest_barycentric(xrt_ctx* ctx, int32_t inst, int32_t* idx, float* uvw) {
  int rc = CheckInstance(ctx, inst, true);
  if (rc != XRT_OK) return rc;
  XRT_REQUIRE(idx && uvw, XRT_ERR_INVALID_ARG, "null output");
  ContactData* c = ctx->contact[inst].get();
  XRT_REQUIRE(c && c->obs_nf >= 2, XRT_ERR_STATE, "instance has no obstacle of >= 2 faces");
  BindingData& B = ctx->binding;
  const size_t N = static_cast<size_t>(ctx->wide.n_total);
  if (B.near_idx.n < N) XRT_CHECK_HIP(B.near_idx.Alloc(N));
  if (B.near_uvw.n < N) XRT_CHECK_HIP(B.near_uvw.Alloc(N));
  if (!B.flag.ptr) XRT_CHECK_HIP(B.flag.Alloc(1));
  if ((rc = NearestRecords(ctx, inst)) != XRT_OK) return rc;
  const InstDesc& d = ctx->inst[inst].desc;
  std::vector<int4> hi(d.nv);
  std::vector<float4> hu(d.nv);
  XRT_CHECK_HIP(hipMemcpyAsync(hi.data(), B.near_idx.ptr + d.vbeg, sizeof(int4) * d.nv,
                               hipMemcpyDeviceToHost, ctx->stream));
  XRT_CHECK_HIP(hipMemcpyAsync(hu.data(), B.near_uvw.ptr + d.vbeg, sizeof(float4) * d.nv,
                               hipMemcpyDeviceToHost, ctx->stream));
  XRT_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  for (int v = 0; v < d.nv; ++v) {
    idx[4 * v] = hi[v].x; idx[4 * v + 1] = hi[v].y; idx[4 * v + 2] = hi[v].z; idx[4 * v + 3] = hi[v].w;
    uvw[4 * v] = hu[v].x; uvw[4 * v + 1] = hu[v].y; uvw[4 * v + 2] = hu[v].z; uvw[4 * v + 3] = hu[v].w;
  }
  return XRT_OK;
}

int xrt_get_bindings(xrt_ctx* ctx, int32_t inst, int32_t* idx, float* par) {
  int rc = CheckInstance(ctx, inst, true);
  if (rc != XRT_OK) return rc;
  XRT_REQUIRE(idx && par, XRT_ERR_INVALID_ARG, "null output");
  const BindInstDev& bi = ctx->binding.inst[inst];
  std::vector<BindItem> it(std::max(bi.nb, 1));
  if (bi.nb > 0) {
    XRT_CHECK_HIP(hipMemcpyAsync(it.data(), ctx->binding.items.ptr + bi.bbeg, sizeof(BindItem) * bi.nb,
                                 hipMemcpyDeviceToHost, ctx->stream));
    XRT_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  }
  for (int b = 0; b < bi.nb; ++b) {
    idx[4 * b] = it[b].idx.x; idx[4 * b + 1] = it[b].idx.y;
    idx[4 * b + 2] = it[b].idx.z; idx[4 * b + 3] = it[b].idx.w;
    par[5 * b] = it[b].uvwl.x; par[5 * b + 1] = it[b].uvwl.y; par[5 * b + 2] = it[b].uvwl.z;
    par[5 * b + 3] = it[b].uvwl.w; par[5 * b + 4] = it[b].k;
  }
  return XRT_OK;
}

int xrt_solve_bindings(xrt_ctx* ctx, int32_t inst) {
  int rc = CheckInstance(ctx, inst, true);
  if (rc != XRT_OK) return rc;
  return BindingIter(ctx, inst);
}

int xrt_time_binding(xrt_ctx* ctx, int32_t n, float ms[2]) {
  XRT_REQUIRE(ctx && ctx->started, XRT_ERR_STATE, "xrt_start has not been called");
  XRT_REQUIRE(n > 0 && ms, XRT_ERR_INVALID_ARG, "n must be > 0, output non-null");
  BindingData& B = ctx->binding;
  hipStream_t s = ctx->stream;
  std::vector<hipEvent_t> ev(4 * static_cast<size_t>(n), nullptr);
  int rc = XRT_OK;
  for (auto& e : ev)
    if (hipEventCreate(&e) != hipSuccess) rc = XRT_ERR_HIP;
  for (int k = 0; k < n && rc == XRT_OK; ++k) {
    if (hipEventRecord(ev[4 * k], s) != hipSuccess ||
        LaunchLbs(s, B.skin_dev.ptr, B.transforms.ptr, Map(B.skin_maps, B.skin_all)) != hipSuccess ||
        hipEventRecord(ev[4 * k + 1], s) != hipSuccess || hipEventRecord(ev[4 * k + 2], s) != hipSuccess)
      rc = XRT_ERR_HIP;
    if (rc == XRT_OK && BindingIter(ctx, -1) != XRT_OK) rc = XRT_ERR_HIP;
    if (rc == XRT_OK && hipEventRecord(ev[4 * k + 3], s) != hipSuccess) rc = XRT_ERR_HIP;
  }
  if (rc == XRT_OK && hipStreamSynchronize(s) != hipSuccess) rc = XRT_ERR_HIP;
  float tot[2] = {0.0f, 0.0f};
  for (int k = 0; k < n && rc == XRT_OK; ++k)
    for (int j = 0; j < 2; ++j) {
      float t = 0.0f;
      if (hipEventElapsedTime(&t, ev[4 * k + 2 * j], ev[4 * k + 2 * j + 1]) != hipSuccess) rc = XRT_ERR_HIP;
      tot[j] += t;
    }
  for (auto& e : ev)
    if (e) (void)hipEventDestroy(e);
  if (rc != XRT_OK) {
    SetError("HIP error while timing the skinning / binding kernels");
    return rc;
  }
  ms[0] = tot[0] / static_cast<float>(n);
  ms[1] = tot[1] / static_cast<float>(n);
  return XRT_OK;
}

int xrt_get_obstacle(xrt_ctx* ctx, int32_t inst, float* x0, float* x) {
  int rc = CheckInstance(ctx, inst, true);
  if (rc != XRT_OK) return rc;
  ContactData* c = ctx->contact[inst].get();
  XRT_REQUIRE(c && c->obs_nf > 0, XRT_ERR_STATE, "instance has no obstacle");
  std::vector<float4> buf(c->obs_nv);
  for (int pass = 0; pass < 2; ++pass) {
    float* out = pass ? x : x0;
    if (!out) continue;
    XRT_CHECK_HIP(hipMemcpyAsync(buf.data(), pass ? c->o_x.ptr : c->o_x0.ptr,
                                 sizeof(float4) * c->obs_nv, hipMemcpyDeviceToHost, ctx->stream));
    XRT_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    for (int v = 0; v < c->obs_nv; ++v) {
      out[3 * v] = buf[v].x; out[3 * v + 1] = buf[v].y; out[3 * v + 2] = buf[v].z;
    }
  }
  return XRT_OK;
}

}  // extern "C"
