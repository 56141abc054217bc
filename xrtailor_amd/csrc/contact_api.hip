// Contact set-up and the contact-aware plain substep.
//
// Substep order (a union of ClothSolverBase::SimulateSwift's body-cloth collision and
// particle self-collision, xrtailor/src/pipeline/base/ClothSolverBase.cpp:190-240, and
// SimulateQuality's PBD repulsion, :300-313):
//   predict; spatial hash when the frame's substep % inter_leaved_hash == 0
//   per iteration:
//     particle collision (CollideParticles + ApplyDeltas x relaxation factor)
//     collision pass (step % frequency == 0): obstacle x1 = lerp(x0, x, (cnt+1)/passes),
//       midstep normals, RefitMidstep(bvh_tolerance); generate VF-DCD, VF-CCD, EE-DCD
//     pair refresh (iteration % pair_update_interval == 0): cloth LBVH Update(ccd),
//       Traverse(cloth, obstacle) ++ Traverse(cloth, cloth) at thickness 1e-6
//     repulsion over the pairs -> impulses x relaxation rate
//     stretch colors, bend colors, LRA, ground
//     solve VF-CCD, VF-DCD (direct), EE-DCD (deltas) -> ApplyDeltas x relaxation factor
//     bindings (BindingConstraint::Solve against the obstacle's frame positions)
//   finalize
// Frame start: obstacle LBVH rebuilt from its x0 (tol 0), cloth LBVH rebuilt from x0.
// Frame end: obstacle x0 = x (PhysicsMesh::PredictedToPositions).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "contact.h"
#include "context.h"
#include "dynamics.h"

namespace xrt {

namespace {
constexpr float kPairThickness = 1e-6f;  // UpdatePairs thickness (ClothSolverBase.cpp:305-308)

int UploadTopo(const ContactTopo& t, const std::vector<int32_t>& tris,
               const std::vector<int32_t>& edges, DevBuf<int4>& dt, DevBuf<int4>& dfe,
               DevBuf<int2>& de, DevBuf<int32_t>& dptr, DevBuf<int32_t>& def,
               DevBuf<float4>& dn, DevBuf<uint32_t>& dr, DevBuf<int4>& dtr, hipStream_t s) {
  const int nt = static_cast<int>(tris.size() / 3), ne = static_cast<int>(edges.size() / 2);
  std::vector<int4> t4(std::max(nt, 1)), fe4(std::max(nt, 1)), tr4(std::max(nt, 1));
  std::vector<float4> n4(std::max(nt, 1));
  for (int f = 0; f < nt; ++f) {
    t4[f] = make_int4(tris[3 * f], tris[3 * f + 1], tris[3 * f + 2], 0);
    tr4[f] = make_int4(tris[3 * f], tris[3 * f + 1], tris[3 * f + 2], static_cast<int>(t.rtri[f]));
    fe4[f] = make_int4(t.face_edges[3 * f], t.face_edges[3 * f + 1], t.face_edges[3 * f + 2], 0);
    n4[f] = make_float4(t.face_n[3 * f], t.face_n[3 * f + 1], t.face_n[3 * f + 2], 0.0f);
  }
  std::vector<int2> e2(std::max(ne, 1));
  for (int e = 0; e < ne; ++e) e2[e] = make_int2(edges[2 * e], edges[2 * e + 1]);
  XRT_CHECK_HIP(dt.Upload(t4, s));
  XRT_CHECK_HIP(dfe.Upload(fe4, s));
  XRT_CHECK_HIP(dn.Upload(n4, s));
  XRT_CHECK_HIP(de.Upload(e2, s));
  XRT_CHECK_HIP(dptr.Upload(t.ef_ptr, s));
  XRT_CHECK_HIP(def.Upload(t.ef, s));
  XRT_CHECK_HIP(dr.Upload(t.rtri, s));
  XRT_CHECK_HIP(dtr.Upload(tr4, s));
  return XRT_OK;
}

MeshTopo ClothTopo(const ContactData& c, const HostInstance& h) {
  MeshTopo t;
  t.nv = h.nv;
  t.nf = h.nt;
  t.ne = h.ne();
  t.tris = c.c_tris.ptr;
  t.edges = c.c_edges.ptr;
  t.face_edges = c.c_fe.ptr;
  t.ef_ptr = c.c_ef_ptr.ptr;
  t.ef = c.c_ef.ptr;
  t.face_n = c.c_face_n.ptr;
  t.rtri = c.c_rtri.ptr;
  t.trir = c.c_trir.ptr;
  return t;
}

MeshTopo ObsTopo(const ContactData& c) {
  MeshTopo t;
  t.nv = c.obs_nv;
  t.nf = c.obs_nf;
  t.ne = c.obs_ne;
  t.tris = c.o_tris.ptr;
  t.edges = c.o_edges.ptr;
  t.face_edges = c.o_fe.ptr;
  t.ef_ptr = c.o_ef_ptr.ptr;
  t.ef = c.o_ef.ptr;
  t.face_n = c.o_face_n.ptr;
  t.rtri = c.o_rtri.ptr;
  t.trir = c.o_trir.ptr;
  return t;
}

// friction of instance inst's VF solves (contact_friction; 0: none)
VfFriction Friction(const xrt_ctx* ctx, int inst) {
  const ContactData& c = *ctx->contact[inst];
  return VfFriction{ctx->x0.ptr + ctx->inst[inst].desc.vbeg, c.o_x0.ptr, c.o_x.ptr,
                    1.0f / static_cast<float>(ctx->params.num_substeps), ctx->params.contact_friction};
}

std::vector<float4> ToF4(const std::vector<float>& p) {
  std::vector<float4> out(std::max<size_t>(p.size() / 3, 1));
  for (size_t i = 0; i < p.size() / 3; ++i) out[i] = make_float4(p[3 * i], p[3 * i + 1], p[3 * i + 2], 0.0f);
  return out;
}
}  // namespace

bool ContactNeeded(const xrt_ctx* ctx, const HostInstance& h) {
  return h.obs_nt() > 0 || ctx->params.enable_self_contact;
}

// pair list and repulsion item lists of one instance at the context's capacities
int AllocPairLists(xrt_ctx* ctx, const HostInstance& h, ContactData& c) {
  const int64_t per_face = std::max(1, ctx->params.pair_capacity_per_face);
  XRT_REQUIRE(per_face * std::max(h.nt, 1) <= static_cast<int64_t>(INT32_MAX), XRT_ERR_INVALID_ARG,
              "pair_capacity_per_face x faces exceeds 2^31 - 1");
  c.pair_cap = static_cast<int>(std::max<int64_t>(1, per_face * h.nt));
  XRT_CHECK_HIP(c.pairs.Alloc(c.pair_cap));
  // repulsion item lists (at most 3 VF and 9 EE items per pair; an instance whose items
  // overflow a list falls back to the per-pair tests for that half); off when either
  // capacity is 0 or the counts could overflow an int
  const int64_t vf_cap = static_cast<int64_t>(std::max(0, ctx->params.repulsion_vf_items_per_face)) * h.nt;
  const int64_t ee_cap = static_cast<int64_t>(std::max(0, ctx->params.repulsion_ee_items_per_face)) * h.nt;
  c.vf_cap = c.ee_cap = 0;
  c.rep_vf.Free();
  c.rep_ee.Free();
  if (ctx->params.enable_self_contact && vf_cap > 0 && ee_cap > 0 &&
      9 * static_cast<int64_t>(c.pair_cap) <= static_cast<int64_t>(INT32_MAX)) {
    c.vf_cap = static_cast<int>(std::min<int64_t>(vf_cap, 3 * static_cast<int64_t>(c.pair_cap)));
    c.ee_cap = static_cast<int>(std::min<int64_t>(ee_cap, 9 * static_cast<int64_t>(c.pair_cap)));
    XRT_CHECK_HIP(c.rep_vf.Alloc(c.vf_cap));
    XRT_CHECK_HIP(c.rep_ee.Alloc(c.ee_cap));
    if (!c.ivf.ptr) XRT_CHECK_HIP(c.ivf.Alloc(std::max(h.nt, 1)));
    if (!c.iee.ptr) XRT_CHECK_HIP(c.iee.Alloc(std::max(h.nt, 1)));
    if (!c.nitem.ptr) XRT_CHECK_HIP(c.nitem.Alloc(std::max(h.nt, 1)));
  }
  return XRT_OK;
}

int ContactStart(xrt_ctx* ctx) {
  hipStream_t s = ctx->stream;
  XRT_CHECK_HIP(ctx->overflow.Alloc(1));
  XRT_CHECK_HIP(hipMemsetAsync(ctx->overflow.ptr, 0, sizeof(int32_t), s));
  ctx->contact.clear();
  ctx->contact.resize(ctx->inst.size());
  for (size_t ii = 0; ii < ctx->inst.size(); ++ii) {
    HostInstance& h = ctx->inst[ii];
    if (!ContactNeeded(ctx, h)) continue;
    auto c = std::make_unique<ContactData>();
    h.cloth_topo = BuildContactTopo(h.rest_pos.data(), h.nv, h.tris.data(), h.nt, h.edges);
    int rc = UploadTopo(h.cloth_topo, h.tris, h.edges, c->c_tris, c->c_fe, c->c_edges, c->c_ef_ptr,
                        c->c_ef, c->c_face_n, c->c_rtri, c->c_trir, s);
    if (rc != XRT_OK) return rc;
    XRT_CHECK_HIP(c->cloth_bvh.Setup(h.tris.data(), h.nt, h.nv, s));
    if (h.obs_nt() > 0) {
      c->obs_nv = h.obs_nv();
      c->obs_nf = h.obs_nt();
      c->obs_ne = static_cast<int>(h.obs_edges.size() / 2);
      h.obs_topo = BuildContactTopo(h.obs_pos.data(), c->obs_nv, h.obs_tris.data(), c->obs_nf,
                                    h.obs_edges);
      rc = UploadTopo(h.obs_topo, h.obs_tris, h.obs_edges, c->o_tris, c->o_fe, c->o_edges,
                      c->o_ef_ptr, c->o_ef, c->o_face_n, c->o_rtri, c->o_trir, s);
      if (rc != XRT_OK) return rc;
      const std::vector<float4> p = ToF4(h.obs_pos);
      XRT_CHECK_HIP(c->o_x0.Upload(p, s));
      XRT_CHECK_HIP(c->o_x.Upload(p, s));
      XRT_CHECK_HIP(c->o_x1.Upload(p, s));
      XRT_CHECK_HIP(c->o_n1.Alloc(p.size()));
      // vertex -> corner CSR (ascending face order) for the midstep normals
      std::vector<int32_t> vptr(c->obs_nv + 1, 0), corner(3 * c->obs_nf);
      for (int i = 0; i < 3 * c->obs_nf; ++i) vptr[h.obs_tris[i] + 1]++;
      for (int v = 0; v < c->obs_nv; ++v) vptr[v + 1] += vptr[v];
      std::vector<int32_t> fill(vptr.begin(), vptr.end() - 1);
      for (int i = 0; i < 3 * c->obs_nf; ++i) corner[fill[h.obs_tris[i]]++] = i;
      XRT_CHECK_HIP(c->o_vptr.Upload(vptr, s));
      XRT_CHECK_HIP(c->o_corner.Upload(corner, s));
      XRT_CHECK_HIP(c->obs_bvh.Setup(h.obs_tris.data(), c->obs_nf, c->obs_nv, s));
    }
    XRT_CHECK_HIP(c->vf_face.Alloc(h.nv));
    XRT_CHECK_HIP(c->rt_face.Alloc(h.nv));
    XRT_CHECK_HIP(c->vf_q.Alloc(h.nv));
    XRT_CHECK_HIP(c->rt_q.Alloc(h.nv));
    XRT_CHECK_HIP(c->ee.Alloc(sizeof(EeRec) * std::max(h.ne(), 1)));
    XRT_CHECK_HIP(hipMemsetAsync(c->ee.ptr, 0, c->ee.n, s));  // no records before a pass
    XRT_CHECK_HIP(c->ee_none.Alloc(std::max(h.ne(), 1)));      // (zeroed records: none = 0)
    XRT_CHECK_HIP(hipMemsetAsync(c->ee_none.ptr, 0, c->ee_none.n, s));
    XRT_CHECK_HIP(c->dq.Alloc(3 * static_cast<size_t>(h.nv)));
    XRT_CHECK_HIP(c->cnt.Alloc(h.nv));
    XRT_CHECK_HIP(hipMemsetAsync(c->dq.ptr, 0, sizeof(long long) * 3 * h.nv, s));
    XRT_CHECK_HIP(hipMemsetAsync(c->cnt.ptr, 0, sizeof(int32_t) * h.nv, s));
    XRT_CHECK_HIP(hipMemsetAsync(c->vf_face.ptr, 0xff, sizeof(int32_t) * h.nv, s));
    XRT_CHECK_HIP(hipMemsetAsync(c->rt_face.ptr, 0xff, sizeof(int32_t) * h.nv, s));
    XRT_CHECK_HIP(c->scal.Alloc(16));
    XRT_CHECK_HIP(hipMemsetAsync(c->scal.ptr, 0, sizeof(int32_t) * 16, s));
    rc = AllocPairLists(ctx, h, *c);
    if (rc != XRT_OK) return rc;
    if (ctx->params.enable_self_contact && h.nt < (1 << kStageObjBits) &&
        c->obs_nf < (1 << kStageObjBits))
      XRT_CHECK_HIP(c->stage.Alloc(static_cast<size_t>(kPairStage) * std::max(h.nt, 1)));
    XRT_CHECK_HIP(c->bsum.Alloc(3 * static_cast<size_t>(std::max(1, (h.nt + 255) / 256))));
    ctx->contact[ii] = std::move(c);
  }
  XRT_CHECK_HIP(hipStreamSynchronize(s));
  return MultiContactSetup(ctx);
}

// A state set from outside (xrt_set_state on a started context) voids what earlier substeps
// derived from the old state: the obstacle contact records (VF / RT / EE) are cleared, so no
// record acts before the instance's next collision pass, and the cloth tree is rebuilt from
// the new positions (a fresh start builds it from the initial state the same way).
int ContactStateReset(xrt_ctx* ctx, int inst) {
  if (inst < 0 || static_cast<size_t>(inst) >= ctx->contact.size() || !ctx->contact[inst])
    return XRT_OK;
  ContactData& c = *ctx->contact[inst];
  const HostInstance& h = ctx->inst[inst];
  hipStream_t s = ctx->stream;
  XRT_CHECK_HIP(hipMemsetAsync(c.vf_face.ptr, 0xff, sizeof(int32_t) * h.nv, s));
  XRT_CHECK_HIP(hipMemsetAsync(c.rt_face.ptr, 0xff, sizeof(int32_t) * h.nv, s));
  XRT_CHECK_HIP(hipMemsetAsync(c.ee.ptr, 0, c.ee.n, s));
  XRT_CHECK_HIP(hipMemsetAsync(c.ee_none.ptr, 0, c.ee_none.n, s));
  if (ctx->params.enable_self_contact)
    XRT_CHECK_HIP(LbvhBuild(c.cloth_bvh, ctx->x0.ptr + h.desc.vbeg, nullptr, nullptr, 0.0f, 0,
                            0.0f, s));
  XRT_CHECK_HIP(hipStreamSynchronize(s));
  return XRT_OK;
}

// Contact phases of one instance's substep, shared by the plain driver below and the
// wide (all instances per launch) driver of the adjoint.
int ContactSubstepBegin(xrt_ctx* ctx, int inst) {
  ContactData& c = *ctx->contact[inst];
  hipStream_t s = ctx->stream;
  if (c.frame_substep == 0) {  // frame start
    if (c.obs_nf > 0)
      XRT_CHECK_HIP(LbvhBuild(c.obs_bvh, c.o_x0.ptr, nullptr, nullptr, 0.0f, 0, 0.0f, s));
    if (ctx->params.enable_self_contact)
      XRT_CHECK_HIP(LbvhBuild(c.cloth_bvh, ctx->x0.ptr + ctx->inst[inst].desc.vbeg, nullptr,
                              nullptr, 0.0f, 0, 0.0f, s));
    c.pass_cnt = 0;
  }
  return XRT_OK;
}

// The phases below are what both the plain driver (ContactIterPre) and the reference-shaped
// entry points (xrt_vf_dcd_generate, xrt_repulsion_update_pairs, ...) launch, so a caller
// that drives the reference's order through the entry points gets the driver's bits.

// obstacle x1 = lerp(x0, x, (cnt + 1) / passes) (PhysicsMesh::Interpolate)
int PhaseObstacleInterpolate(xrt_ctx* ctx, int inst, int cnt, int passes) {
  ContactData& c = *ctx->contact[inst];
  const float alpha = (static_cast<float>(cnt) + 1.0f) / static_cast<float>(passes);
  XRT_CHECK_HIP(LaunchInterpolate(ctx->stream, c.o_x0.ptr, c.o_x.ptr, c.o_x1.ptr, c.obs_nv, alpha));
  return XRT_OK;
}

// n1 = normals of x1 (PhysicsMesh::UpdateMidstepNormals)
int PhaseObstacleMidstepNormals(xrt_ctx* ctx, int inst) {
  ContactData& c = *ctx->contact[inst];
  XRT_CHECK_HIP(VertexNormals(c.o_x1.ptr, c.o_tris.ptr, c.o_vptr.ptr, c.o_corner.ptr, c.obs_nv,
                              c.o_n1.ptr, ctx->stream));
  return XRT_OK;
}

// BVH::RefitMidstep(faces, tolerance) of the obstacle tree
int PhaseObstacleRefitMidstep(xrt_ctx* ctx, int inst, float tol) {
  ContactData& c = *ctx->contact[inst];
  XRT_CHECK_HIP(LbvhRefitMidstep(c.obs_bvh, c.o_x1.ptr, c.o_n1.ptr, tol, ctx->stream));
  return XRT_OK;
}

// kind 0: VF-DCD, 1: VF-CCD (RTConstraint), 2: EE-DCD records against the obstacle tree
int PhaseGenerate(xrt_ctx* ctx, int inst, int kind) {
  const HostInstance& h = ctx->inst[inst];
  ContactData& c = *ctx->contact[inst];
  hipStream_t s = ctx->stream;
  float4* x0 = ctx->x0.ptr + h.desc.vbeg;
  float4* x = ctx->x.ptr + h.desc.vbeg;
  if (kind == 0) XRT_CHECK_HIP(LaunchVfDcdGenerate(s, x, h.nv, c.obs_bvh, c.vf_face.ptr, c.vf_q.ptr));
  if (kind == 1) XRT_CHECK_HIP(LaunchVfCcdGenerate(s, x0, x, h.nv, c.obs_bvh, c.rt_face.ptr, c.rt_q.ptr));
  if (kind == 2)
    XRT_CHECK_HIP(LaunchEeDcdGenerate(s, x, c.c_edges.ptr, h.ne(), c.obs_bvh, c.o_rtri.ptr,
                                      reinterpret_cast<EeRec*>(c.ee.ptr), c.ee_none.ptr));
  return XRT_OK;
}

// PBDRepulsionSolver::UpdatePairs: cloth LBVH Update(x, x0, ccd), Traverse(cloth, obstacle)
// at [0, n_obs), Traverse(cloth, cloth) after them
int PhaseUpdatePairs(xrt_ctx* ctx, int inst, float thickness) {
  const HostInstance& h = ctx->inst[inst];
  ContactData& c = *ctx->contact[inst];
  hipStream_t s = ctx->stream;
  float4* x0 = ctx->x0.ptr + h.desc.vbeg;
  float4* x = ctx->x.ptr + h.desc.vbeg;
  XRT_CHECK_HIP(LbvhUpdate(c.cloth_bvh, x, x0, 1, s));
  if (c.obs_nf > 0) {
    XRT_CHECK_HIP(LbvhCountPairs(c.cloth_bvh, c.obs_bvh, thickness, s));
    XRT_CHECK_HIP(hipMemcpyAsync(c.scal.ptr + 1, c.cloth_bvh.total.ptr, 4, hipMemcpyDeviceToDevice, s));
    XRT_CHECK_HIP(LbvhFindPairsAt(c.cloth_bvh, c.obs_bvh, thickness, c.pairs.ptr, c.scal.ptr,
                                  c.pair_cap, s));
  } else {
    XRT_CHECK_HIP(hipMemsetAsync(c.scal.ptr + 1, 0, 4, s));
  }
  XRT_CHECK_HIP(LbvhCountPairs(c.cloth_bvh, c.cloth_bvh, thickness, s));
  XRT_CHECK_HIP(LbvhFindPairsAt(c.cloth_bvh, c.cloth_bvh, thickness, c.pairs.ptr, c.scal.ptr + 1,
                                c.pair_cap, s));
  XRT_CHECK_HIP(LaunchAddTotals(s, c.scal.ptr + 1, c.cloth_bvh.total.ptr, c.scal.ptr + 2,
                                c.pair_cap, h.nt, ctx->overflow.ptr));
  return XRT_OK;
}

// PBDRepulsionSolver::SolveNoProximity: per-pair VF / EE tests -> Jacobi sums -> x relax
int PhaseRepulsion(xrt_ctx* ctx, int inst, float thickness, float relax) {
  const HostInstance& h = ctx->inst[inst];
  ContactData& c = *ctx->contact[inst];
  hipStream_t s = ctx->stream;
  float4* x0 = ctx->x0.ptr + h.desc.vbeg;
  float4* x = ctx->x.ptr + h.desc.vbeg;
  MeshState cs{x, x0, -1.0f};
  MeshState os{c.o_x.ptr, c.o_x0.ptr, 1e-3f};  // obstacle nodes: inv mass 1e-3, not free
  XRT_CHECK_HIP(LaunchRepulsion(s, ClothTopo(c, h), cs, ObsTopo(c), os, c.pairs.ptr, c.scal.ptr + 1,
                                c.scal.ptr + 2, c.pair_cap, thickness, c.dq.ptr, c.cnt.ptr));
  XRT_CHECK_HIP(LaunchApplyFixedDeltas(s, x, h.nv, c.dq.ptr, c.cnt.ptr, relax));
  return XRT_OK;
}

// collision pass when due (*pass_ran says whether it ran), pair refresh and repulsion
int ContactIterPre(xrt_ctx* ctx, int inst, int it, bool* pass_ran) {
  ContactData& c = *ctx->contact[inst];
  const xrt_params& p = ctx->params;
  const bool has_obs = c.obs_nf > 0 && p.enable_obstacle_contact;
  const bool self = p.enable_self_contact != 0;
  const int frequency = std::max(1, p.num_substeps * p.num_iterations / p.num_collision_passes);
  const int interval = std::max(1, p.pair_update_interval);
  int rc = XRT_OK;
  if (pass_ran) *pass_ran = false;
  const int step = c.frame_substep * p.num_iterations + it;
  if (has_obs && step % frequency == 0) {
    if ((rc = PhaseObstacleInterpolate(ctx, inst, c.pass_cnt, p.num_collision_passes)) != XRT_OK ||
        (rc = PhaseObstacleMidstepNormals(ctx, inst)) != XRT_OK ||
        (rc = PhaseObstacleRefitMidstep(ctx, inst, p.bvh_tolerance)) != XRT_OK)
      return rc;
    for (int kind = 0; kind < 3; ++kind)
      if ((rc = PhaseGenerate(ctx, inst, kind)) != XRT_OK) return rc;
    c.pass_cnt++;
    if (pass_ran) *pass_ran = true;
  }
  if (self && it % interval == 0 && (rc = PhaseUpdatePairs(ctx, inst, kPairThickness)) != XRT_OK)
    return rc;
  if (self) return PhaseRepulsion(ctx, inst, p.pbd_thickness, p.repulsion_relaxation);
  return XRT_OK;
}

// VF-CCD solve (offset 0), then VF-DCD solve (offset 1e-3)
int ContactVfSolve(xrt_ctx* ctx, int inst) {
  ContactData& c = *ctx->contact[inst];
  if (!(c.obs_nf > 0 && ctx->params.enable_obstacle_contact)) return XRT_OK;
  const InstDesc& d = ctx->inst[inst].desc;
  float4* x = ctx->x.ptr + d.vbeg;
  const VfFriction fr = Friction(ctx, inst);
  XRT_CHECK_HIP(LaunchVfSolve(ctx->stream, x, d.nv, c.obs_bvh, c.rt_face.ptr, c.rt_q.ptr, 0.0f, fr));
  XRT_CHECK_HIP(LaunchVfSolve(ctx->stream, x, d.nv, c.obs_bvh, c.vf_face.ptr, c.vf_q.ptr, 1e-3f, fr));
  return XRT_OK;
}

// EE-DCD deltas -> ApplyDeltas x relaxation factor
int ContactEeSolve(xrt_ctx* ctx, int inst) {
  ContactData& c = *ctx->contact[inst];
  if (!(c.obs_nf > 0 && ctx->params.enable_obstacle_contact)) return XRT_OK;
  const HostInstance& h = ctx->inst[inst];
  float4* x = ctx->x.ptr + h.desc.vbeg;
  EeRec* ee = reinterpret_cast<EeRec*>(c.ee.ptr);
  XRT_CHECK_HIP(LaunchEeSolve(ctx->stream, x, c.c_edges.ptr, h.ne(), ee, c.ee_none.ptr, c.dq.ptr,
                              c.cnt.ptr));
  XRT_CHECK_HIP(LaunchApplyFixedDeltas(ctx->stream, x, h.desc.nv, c.dq.ptr, c.cnt.ptr,
                                       ctx->params.relaxation_factor));
  return XRT_OK;
}

// frame bookkeeping after finalize; frame end: obstacle x0 = x
int ContactSubstepEnd(xrt_ctx* ctx, int inst) {
  ContactData& c = *ctx->contact[inst];
  if (++c.frame_substep >= ctx->params.num_substeps) {
    c.frame_substep = 0;
    c.obstacle_moving = false;
    if (c.obs_nf > 0)
      XRT_CHECK_HIP(hipMemcpyAsync(c.o_x0.ptr, c.o_x.ptr, sizeof(float4) * c.obs_nv,
                                   hipMemcpyDeviceToDevice, ctx->stream));
  }
  return XRT_OK;
}

// ------------------------------------------------------------ every instance at once
bool ContactsActive(const xrt_ctx* ctx) {
  if (ctx->params.enable_self_contact) return true;
  if (!ctx->params.enable_obstacle_contact) return false;
  for (const auto& c : ctx->contact)
    if (c && c->obs_nf > 0) return true;
  return false;
}

// device views of every instance's contact state and the block -> instance maps
int ContactResizeLists(xrt_ctx* ctx) {
  XRT_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  for (size_t k = 0; k < ctx->inst.size(); ++k) {
    if (!ctx->contact[k]) continue;
    const int rc = AllocPairLists(ctx, ctx->inst[k], *ctx->contact[k]);
    if (rc != XRT_OK) return rc;
  }
  return MultiContactSetup(ctx);
}

int CheckPairOverflow(xrt_ctx* ctx) {
  XRT_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  if (!ctx->overflow.ptr) return XRT_OK;
  int32_t need = 0;
  XRT_CHECK_HIP(hipMemcpy(&need, ctx->overflow.ptr, sizeof(need), hipMemcpyDeviceToHost));
  if (need <= 0) return XRT_OK;
  XRT_CHECK_HIP(hipMemsetAsync(ctx->overflow.ptr, 0, sizeof(int32_t), ctx->stream));
  XRT_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  SetError("face-pair list overflow: an instance needed " + std::to_string(need) +
           " pairs per face, pair_capacity_per_face is " +
           std::to_string(ctx->params.pair_capacity_per_face) +
           "; the pairs beyond the capacity were not evaluated. Raise pair_capacity_per_face "
           "(xrt_set_params re-allocates the lists) and run the frames again");
  return XRT_ERR_CAPACITY;
}

int MultiContactSetup(xrt_ctx* ctx) {
  const int n = static_cast<int>(ctx->inst.size());
  std::vector<InstContactDev> ic(std::max(n, 1));
  std::vector<int32_t> off(MultiMaps::kKinds * static_cast<size_t>(n + 1), 0);
  auto blocks = [](int count) { return (count + 255) / 256; };
  for (int k = 0; k < n; ++k) {
    int cnt[MultiMaps::kKinds] = {0, 0, 0, 0, 0, 0};
    if (ctx->contact[k]) {
      HostInstance& h = ctx->inst[k];
      ContactData& c = *ctx->contact[k];
      InstContactDev& d = ic[k];
      d.vbeg = h.desc.vbeg;
      d.nv = h.nv;
      d.nf = h.nt;
      d.ne = h.ne();
      d.has_obs = c.obs_nf > 0;
      d.obs_nv = c.obs_nv;
      d.obs_nf = c.obs_nf;
      d.pair_cap = c.pair_cap;
      d.cloth = ClothTopo(c, h);
      d.obs = ObsTopo(c);
      d.cbvh = c.cloth_bvh.view();
      d.obvh = c.obs_bvh.view();
      d.o_x0 = c.o_x0.ptr;
      d.o_x = c.o_x.ptr;
      d.o_x1 = c.o_x1.ptr;
      d.o_n1 = c.o_n1.ptr;
      d.o_vptr = c.o_vptr.ptr;
      d.o_corner = c.o_corner.ptr;
      d.vf_face = c.vf_face.ptr;
      d.rt_face = c.rt_face.ptr;
      d.vf_q = c.vf_q.ptr;
      d.rt_q = c.rt_q.ptr;
      d.ee = reinterpret_cast<EeRec*>(c.ee.ptr);
      d.ee_none = c.ee_none.ptr;
      d.dq = c.dq.ptr;
      d.cnt = c.cnt.ptr;
      d.pairs = c.pairs.ptr;
      d.scal = c.scal.ptr;
      d.vf_cap = c.vf_cap;
      d.ee_cap = c.ee_cap;
      d.rep_vf = c.rep_vf.ptr;
      d.rep_ee = c.rep_ee.ptr;
      d.stage = c.stage.ptr;
      d.bsum = c.bsum.ptr;
      d.overflow = ctx->overflow.ptr;
      d.ivf = c.ivf.ptr;
      d.iee = c.iee.ptr;
      d.nitem = c.nitem.ptr;
      cnt[MultiMaps::kVert] = blocks(h.nv);
      cnt[MultiMaps::kEdge] = blocks(h.ne());
      cnt[MultiMaps::kFace] = blocks(h.nt);
      cnt[MultiMaps::kObsVert] = blocks(c.obs_nv);
      cnt[MultiMaps::kObsFace] = blocks(c.obs_nf);
      // repulsion: about 4 pairs per thread at the usual ~16 pairs per face
      cnt[MultiMaps::kPair] = std::max(1, (16 * h.nt + 1023) / 1024);
    }
    for (int kind = 0; kind < MultiMaps::kKinds; ++kind) {
      int32_t* o = off.data() + static_cast<size_t>(kind) * (n + 1);
      o[k + 1] = o[k] + cnt[kind];
    }
  }
  // device layout per kind: the n + 1 prefix sums, then the instance of every block
  std::vector<int32_t> maps;
  size_t seg[MultiMaps::kKinds];
  for (int kind = 0; kind < MultiMaps::kKinds; ++kind) {
    const int32_t* o = off.data() + static_cast<size_t>(kind) * (n + 1);
    seg[kind] = maps.size();
    maps.insert(maps.end(), o, o + n + 1);
    for (int k = 0; k < n; ++k) maps.insert(maps.end(), static_cast<size_t>(o[k + 1] - o[k]), k);
  }
  XRT_CHECK_HIP(ctx->ic.Upload(ic, ctx->stream));
  XRT_CHECK_HIP(ctx->ic_off.Upload(maps, ctx->stream));
  {  // the trees rebuilt at every frame start of the batched path
    std::vector<Lbvh*> ct, ot;
    std::vector<const float4*> cx, ox;
    for (int k = 0; k < n; ++k) {
      ContactData* c = ctx->contact[k].get();
      if (!c) continue;
      ct.push_back(&c->cloth_bvh);
      cx.push_back(ctx->x0.ptr + ctx->inst[k].desc.vbeg);
      if (c->obs_nf > 0) {
        ot.push_back(&c->obs_bvh);
        ox.push_back(c->o_x0.ptr);
      }
    }
    XRT_CHECK_HIP(ctx->cloth_trees.Setup(ct, cx, ctx->stream));
    XRT_CHECK_HIP(ctx->obs_trees.Setup(ot, ox, ctx->stream));
  }
  ctx->maps.n = n;
  for (int kind = 0; kind < MultiMaps::kKinds; ++kind) {
    ctx->maps.off[kind] = ctx->ic_off.ptr + seg[kind];
    ctx->maps.blocks[kind] = off[static_cast<size_t>(kind) * (n + 1) + n];
  }
  XRT_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  return XRT_OK;
}

namespace {
ContactData* FirstContact(xrt_ctx* ctx) {
  for (auto& c : ctx->contact)
    if (c) return c.get();
  return nullptr;
}
}  // namespace

// frame start: the obstacle trees (and the cloth trees with self contact) of every
// instance are rebuilt together, each build step one launch over all trees, then merged
// in one launch (the same steps as LbvhBuild per tree)
int MultiContactSubstepBegin(xrt_ctx* ctx) {
  ContactData* first = FirstContact(ctx);
  if (!first || first->frame_substep != 0) return XRT_OK;
  hipStream_t s = ctx->stream;
  const bool self = ctx->params.enable_self_contact != 0;
  XRT_CHECK_HIP(MultiLbvhBuild(ctx->obs_trees, s));
  if (self) XRT_CHECK_HIP(MultiLbvhBuild(ctx->cloth_trees, s));
  for (auto& c : ctx->contact)
    if (c) c->pass_cnt = 0;
  XRT_CHECK_HIP(MultiBuildMerge(s, ctx->ic.ptr, ctx->maps, self));
  return XRT_OK;
}

int MultiContactIterPre(xrt_ctx* ctx, int it, bool* pass_ran, const RepTapeHook* rep_tape) {
  if (pass_ran) *pass_ran = false;
  ContactData* first = FirstContact(ctx);
  if (!first) return XRT_OK;
  const xrt_params& p = ctx->params;
  hipStream_t s = ctx->stream;
  const int frequency = std::max(1, p.num_substeps * p.num_iterations / p.num_collision_passes);
  const int interval = std::max(1, p.pair_update_interval);
  const bool obs = ContactsActive(ctx) && p.enable_obstacle_contact &&
                   std::any_of(ctx->contact.begin(), ctx->contact.end(),
                               [](const auto& c) { return c && c->obs_nf > 0; });
  const int step = first->frame_substep * p.num_iterations + it;
  if (obs && step % frequency == 0) {
    int pass = -1;
    for (auto& c : ctx->contact)
      if (c && c->obs_nf > 0) pass = c->pass_cnt;
    const float alpha = (static_cast<float>(pass) + 1.0f) / static_cast<float>(p.num_collision_passes);
    XRT_CHECK_HIP(MultiObstacleRefit(s, ctx->ic.ptr, ctx->maps, alpha, p.bvh_tolerance));
    XRT_CHECK_HIP(MultiGenerate(s, ctx->ic.ptr, ctx->maps, ctx->x0.ptr, ctx->x.ptr));
    for (auto& c : ctx->contact)
      if (c && c->obs_nf > 0) c->pass_cnt++;
    if (pass_ran) *pass_ran = true;
  }
  if (p.enable_self_contact) {
    if (it % interval == 0)
      XRT_CHECK_HIP(MultiPairs(s, ctx->ic.ptr, ctx->maps, ctx->x.ptr, ctx->x0.ptr, kPairThickness));
    if (rep_tape) {
      XRT_CHECK_HIP(hipMemcpyAsync(rep_tape->x_out, ctx->x.ptr, sizeof(float4) * ctx->x.n,
                                   hipMemcpyDeviceToDevice, s));
      XRT_CHECK_HIP(MultiRepRecord(s, ctx->ic.ptr, ctx->maps, ctx->x.ptr, ctx->x0.ptr,
                                   p.pbd_thickness, rep_tape->pool, rep_tape->count,
                                   rep_tape->cap, rep_tape->flags, rep_tape->marks,
                                   rep_tape->marks + 1));
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (ctx->rep_next + 1 < ctx->rep_events.size()) {
      e0 = ctx->rep_events[ctx->rep_next];
      e1 = ctx->rep_events[ctx->rep_next + 1];
      ctx->rep_next += 2;
    }
    XRT_CHECK_HIP(MultiRepulsion(s, ctx->ic.ptr, ctx->maps, ctx->x.ptr, ctx->x0.ptr, p.pbd_thickness,
                                 p.repulsion_relaxation, e0, e1));
  }
  return XRT_OK;
}

int MultiContactVfSolve(xrt_ctx* ctx) {
  if (!ctx->params.enable_obstacle_contact) return XRT_OK;
  XRT_CHECK_HIP(MultiVfSolve(ctx->stream, ctx->ic.ptr, ctx->maps, ctx->x.ptr, ctx->x0.ptr,
                             ctx->params.contact_friction,
                             1.0f / static_cast<float>(ctx->params.num_substeps)));
  return XRT_OK;
}

int MultiContactEeSolve(xrt_ctx* ctx) {
  if (!ctx->params.enable_obstacle_contact) return XRT_OK;
  XRT_CHECK_HIP(MultiEeSolve(ctx->stream, ctx->ic.ptr, ctx->maps, ctx->x.ptr,
                             ctx->params.relaxation_factor));
  return XRT_OK;
}

int MultiContactSubstepEnd(xrt_ctx* ctx) {
  bool frame_end = false;
  for (auto& c : ctx->contact) {
    if (!c) continue;
    if (++c->frame_substep >= ctx->params.num_substeps) {
      c->frame_substep = 0;
      c->obstacle_moving = false;
      frame_end = true;
    }
  }
  if (frame_end) XRT_CHECK_HIP(MultiObstacleFrameEnd(ctx->stream, ctx->ic.ptr, ctx->maps));
  return XRT_OK;
}

// One plain substep of an instance with contacts (see the order at the top of the file).
int SubstepPlainContact(xrt_ctx* ctx, int inst, const StepConsts& k, const BendMode& bm) {
  HostInstance& h = ctx->inst[inst];
  const InstDesc& d = h.desc;
  hipStream_t s = ctx->stream;
  float4* x0 = ctx->x0.ptr + d.vbeg;
  float4* v = ctx->v.ptr + d.vbeg;
  float4* x = ctx->x.ptr + d.vbeg;
  int rc = ContactSubstepBegin(ctx, inst);
  if (rc != XRT_OK) return rc;
  XRT_CHECK_HIP(LaunchPredict(s, x0, v, x, d.nv, k));
  if ((rc = ParticleSubstepBegin(ctx, inst)) != XRT_OK) return rc;
  for (int it = 0; it < k.iterations; ++it) {
    if ((rc = ParticleIter(ctx, inst)) != XRT_OK) return rc;
    if ((rc = ContactIterPre(ctx, inst, it, nullptr)) != XRT_OK) return rc;
    if (FemOn(ctx)) {
      if ((rc = FemIter(ctx, inst, it)) != XRT_OK) return rc;
    } else {
      for (int col = 0; col < h.s_ncol; ++col) {
        const int b = h.sbeg + h.s_coloff[col];
        XRT_CHECK_HIP(LaunchStretch(s, ctx->sc.ptr + b, ctx->slam.ptr + b, x,
                                    h.s_coloff[col + 1] - h.s_coloff[col], k.inv_dt2, it));
      }
    }
    if (k.solve_bending && !FemOn(ctx))
      for (int col = 0; col < h.b_ncol; ++col) {
        const int b = h.bbeg + h.b_coloff[col];
        XRT_CHECK_HIP(LaunchBend(s, ctx->bidx.ptr + b, ctx->bpar.ptr + b, x,
                                 h.b_coloff[col + 1] - h.b_coloff[col], k.inv_dt2, bm));
      }
    if (d.nslots > 0)
      XRT_CHECK_HIP(LaunchLra(s, x, ctx->slot_vert.ptr + d.slot_beg, ctx->slot_pos.ptr, d.nslots,
                              ctx->lra_ptr.ptr + d.lra_vbeg, ctx->lra_slot.ptr, ctx->lra_rest.ptr,
                              d.nv, k.stretchiness));
    if (k.enable_ground) XRT_CHECK_HIP(LaunchGround(s, x, x0, d.nv, k.ground_y, k.friction));
    if ((rc = ContactVfSolve(ctx, inst)) != XRT_OK) return rc;
    if ((rc = ContactEeSolve(ctx, inst)) != XRT_OK) return rc;
    if ((rc = BindingIter(ctx, inst)) != XRT_OK) return rc;
  }
  XRT_CHECK_HIP(LaunchFinalize(s, x0, v, x, d.nv, k));
  if ((rc = ParticleSubstepEnd(ctx, inst)) != XRT_OK) return rc;
  return ContactSubstepEnd(ctx, inst);
}

// ApplyDeltas of one instance's Jacobi sums x relaxation factor (sums cleared)
int ApplyInstanceDeltas(xrt_ctx* ctx, int32_t inst) {
  if (!ctx->contact[inst]) return XRT_OK;  // no contact state: nothing was accumulated
  ContactData& c = *ctx->contact[inst];
  const InstDesc& d = ctx->inst[inst].desc;
  XRT_CHECK_HIP(LaunchApplyFixedDeltas(ctx->stream, ctx->x.ptr + d.vbeg, d.nv, c.dq.ptr, c.cnt.ptr,
                                       ctx->params.relaxation_factor));
  return XRT_OK;
}

}  // namespace xrt

using namespace xrt;

extern "C" {

int xrt_time_repulsion(xrt_ctx* ctx, int32_t n, float* avg_ms, int64_t counts[4]) {
  XRT_REQUIRE(ctx && ctx->started, XRT_ERR_STATE, "xrt_start has not been called");
  XRT_REQUIRE(n > 0 && avg_ms && counts, XRT_ERR_INVALID_ARG, "n must be > 0, outputs non-null");
  XRT_REQUIRE(ctx->params.enable_self_contact && ContactsActive(ctx), XRT_ERR_STATE,
              "the repulsion runs with self contact only");
  XRT_REQUIRE(ctx->rep_events.empty(), XRT_ERR_STATE, "repulsion timing already running");
  hipStream_t s = ctx->stream;
  const size_t pairs = static_cast<size_t>(n) * std::max(1, ctx->params.num_iterations);
  ctx->rep_events.assign(2 * pairs, nullptr);
  ctx->rep_next = 0;
  int rc = XRT_OK;
  for (auto& e : ctx->rep_events)
    if (hipEventCreate(&e) != hipSuccess) rc = XRT_ERR_HIP;
  int64_t sums[4] = {0, 0, 0, 0};
  for (int sub = 0; sub < n && rc == XRT_OK; ++sub) {
    rc = xrt_substep_batched(ctx);
    if (rc != XRT_OK) break;
    // item counts of this substep's last pair refresh
    for (auto& c : ctx->contact) {
      if (!c) continue;
      int32_t sc[9];
      if (hipMemcpyAsync(sc, c->scal.ptr, sizeof(sc), hipMemcpyDeviceToHost, s) != hipSuccess ||
          hipStreamSynchronize(s) != hipSuccess) {
        rc = XRT_ERR_HIP;
        break;
      }
      sums[0] += sc[4];
      sums[1] += sc[3];
      sums[2] += std::min(sc[2], c->pair_cap);
      sums[3] += (c->ee_cap > 0 && sc[2] <= c->pair_cap && sc[4] <= c->ee_cap) ? 1 : 0;
    }
  }
  float total = 0.0f;
  const size_t used = ctx->rep_next / 2;
  if (rc == XRT_OK && hipStreamSynchronize(s) != hipSuccess) rc = XRT_ERR_HIP;
  for (size_t q = 0; q < used && rc == XRT_OK; ++q) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, ctx->rep_events[2 * q], ctx->rep_events[2 * q + 1]) != hipSuccess)
      rc = XRT_ERR_HIP;
    total += ms;
  }
  for (auto& e : ctx->rep_events)
    if (e) (void)hipEventDestroy(e);
  ctx->rep_events.clear();
  ctx->rep_next = 0;
  if (rc != XRT_OK) {
    if (rc == XRT_ERR_HIP) SetError("HIP error while timing the repulsion");
    return rc;
  }
  XRT_REQUIRE(used > 0, XRT_ERR_STATE, "no repulsion launch was timed");
  for (int q = 0; q < 4; ++q) counts[q] = sums[q] / n;
  *avg_ms = total / static_cast<float>(used);
  return XRT_OK;
}

int xrt_set_obstacle(xrt_ctx* ctx, int32_t inst, const float* pos, int32_t nv, const int32_t* tris,
                     int32_t nt) {
  XRT_REQUIRE(ctx && inst >= 0 && inst < static_cast<int32_t>(ctx->inst.size()), XRT_ERR_INVALID_ARG,
              "bad context or instance");
  XRT_REQUIRE(!ctx->started, XRT_ERR_STATE, "set the obstacle before xrt_start");
  XRT_REQUIRE(pos && tris && nv > 0 && nt > 0, XRT_ERR_INVALID_ARG, "empty or null obstacle mesh");
  for (int i = 0; i < 3 * nt; ++i)
    XRT_REQUIRE(tris[i] >= 0 && tris[i] < nv, XRT_ERR_INVALID_ARG, "obstacle face index out of range");
  for (int i = 0; i < 3 * nv; ++i)
    XRT_REQUIRE(std::isfinite(pos[i]), XRT_ERR_NOT_FINITE, "non-finite obstacle position");
  HostInstance& h = ctx->inst[inst];
  h.obs_pos.assign(pos, pos + 3 * nv);
  h.obs_tris.assign(tris, tris + 3 * nt);
  h.obs_edges = BuildEdges(tris, nt);
  return XRT_OK;
}

int xrt_update_obstacle(xrt_ctx* ctx, int32_t inst, const float* pos) {
  XRT_REQUIRE(ctx && inst >= 0 && inst < static_cast<int32_t>(ctx->inst.size()) && pos,
              XRT_ERR_INVALID_ARG, "bad argument");
  XRT_REQUIRE(ctx->started, XRT_ERR_STATE, "xrt_start has not been called");
  ContactData* c = ctx->contact[inst].get();
  XRT_REQUIRE(c && c->obs_nf > 0, XRT_ERR_STATE, "instance has no obstacle");
  std::vector<float> p(pos, pos + 3 * c->obs_nv);
  for (float f : p) XRT_REQUIRE(std::isfinite(f), XRT_ERR_NOT_FINITE, "non-finite obstacle position");
  const std::vector<float4> q = ToF4(p);
  XRT_CHECK_HIP(c->o_x.Upload(q.data(), q.size(), ctx->stream));
  c->obstacle_moving = true;
  XRT_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  return XRT_OK;
}

int xrt_get_contacts(xrt_ctx* ctx, int32_t inst, int32_t* vf_face, float* vf_q, int32_t* rt_face,
                     float* rt_q, float* ee16, int32_t stats[4]) {
  XRT_REQUIRE(ctx && inst >= 0 && inst < static_cast<int32_t>(ctx->inst.size()), XRT_ERR_INVALID_ARG,
              "bad context or instance");
  XRT_REQUIRE(ctx->started, XRT_ERR_STATE, "xrt_start has not been called");
  ContactData* c = ctx->contact[inst].get();
  XRT_REQUIRE(c, XRT_ERR_STATE, "instance has no contact state");
  const HostInstance& h = ctx->inst[inst];
  hipStream_t s = ctx->stream;
  std::vector<float4> q(h.nv);
  if (vf_face) XRT_CHECK_HIP(hipMemcpyAsync(vf_face, c->vf_face.ptr, 4 * h.nv, hipMemcpyDeviceToHost, s));
  if (rt_face) XRT_CHECK_HIP(hipMemcpyAsync(rt_face, c->rt_face.ptr, 4 * h.nv, hipMemcpyDeviceToHost, s));
  if (ee16) XRT_CHECK_HIP(hipMemcpyAsync(ee16, c->ee.ptr, sizeof(EeRec) * h.ne(), hipMemcpyDeviceToHost, s));
  if (stats) {
    int32_t sc[4];
    XRT_CHECK_HIP(hipMemcpyAsync(sc, c->scal.ptr, 16, hipMemcpyDeviceToHost, s));
    XRT_CHECK_HIP(hipStreamSynchronize(s));
    stats[0] = sc[1];
    stats[1] = sc[2];
    stats[2] = c->pair_cap;
    stats[3] = sc[2] > c->pair_cap ? 1 : 0;
  }
  for (int which = 0; which < 2; ++which) {
    float* out = which == 0 ? vf_q : rt_q;
    if (!out) continue;
    XRT_CHECK_HIP(hipMemcpyAsync(q.data(), which == 0 ? c->vf_q.ptr : c->rt_q.ptr, sizeof(float4) * h.nv,
                                 hipMemcpyDeviceToHost, s));
    XRT_CHECK_HIP(hipStreamSynchronize(s));
    for (int i = 0; i < h.nv; ++i) {
      out[3 * i] = q[i].x; out[3 * i + 1] = q[i].y; out[3 * i + 2] = q[i].z;
    }
  }
  XRT_CHECK_HIP(hipStreamSynchronize(s));
  return XRT_OK;
}


// ------------------------------------------------ reference-shaped collider entry points
// One call per reference method, on one instance's contact state; each launches exactly the
// kernels the plain driver launches for that step, so a caller that keeps its own
// SimulateSwift / SimulateQuality loop reproduces xrt_substep_plain bit for bit.
namespace {
int ColliderInst(xrt_ctx* ctx, int32_t inst, bool need_obstacle) {
  XRT_REQUIRE(ctx, XRT_ERR_INVALID_ARG, "null context");
  XRT_REQUIRE(inst >= 0 && inst < static_cast<int32_t>(ctx->inst.size()), XRT_ERR_INVALID_ARG,
              "instance out of range");
  XRT_REQUIRE(ctx->started, XRT_ERR_STATE, "xrt_start has not been called");
  XRT_REQUIRE(ctx->contact[inst] != nullptr, XRT_ERR_STATE,
              "instance has no contact state (no obstacle and self contact off at xrt_start)");
  XRT_REQUIRE(!need_obstacle || ctx->contact[inst]->obs_nf > 0, XRT_ERR_STATE,
              "instance has no obstacle mesh");
  return XRT_OK;
}
}  // namespace

int xrt_obstacle_bvh_construct(xrt_ctx* ctx, int32_t inst) {
  int rc = ColliderInst(ctx, inst, true);
  if (rc != XRT_OK) return rc;
  ContactData& c = *ctx->contact[inst];
  XRT_CHECK_HIP(LbvhBuild(c.obs_bvh, c.o_x0.ptr, nullptr, nullptr, 0.0f, 0, 0.0f, ctx->stream));
  return XRT_OK;
}

int xrt_cloth_bvh_construct(xrt_ctx* ctx, int32_t inst) {
  int rc = ColliderInst(ctx, inst, false);
  if (rc != XRT_OK) return rc;
  XRT_CHECK_HIP(LbvhBuild(ctx->contact[inst]->cloth_bvh, ctx->x0.ptr + ctx->inst[inst].desc.vbeg,
                          nullptr, nullptr, 0.0f, 0, 0.0f, ctx->stream));
  return XRT_OK;
}

int xrt_obstacle_interpolate(xrt_ctx* ctx, int32_t inst, int32_t cnt, int32_t num_passes) {
  int rc = ColliderInst(ctx, inst, true);
  if (rc != XRT_OK) return rc;
  XRT_REQUIRE(num_passes > 0 && cnt >= 0, XRT_ERR_INVALID_ARG, "num_passes must be > 0, cnt >= 0");
  return PhaseObstacleInterpolate(ctx, inst, cnt, num_passes);
}

int xrt_obstacle_update_midstep_normals(xrt_ctx* ctx, int32_t inst) {
  int rc = ColliderInst(ctx, inst, true);
  if (rc != XRT_OK) return rc;
  return PhaseObstacleMidstepNormals(ctx, inst);
}

int xrt_obstacle_bvh_refit_midstep(xrt_ctx* ctx, int32_t inst, float tolerance) {
  int rc = ColliderInst(ctx, inst, true);
  if (rc != XRT_OK) return rc;
  XRT_REQUIRE(std::isfinite(tolerance), XRT_ERR_INVALID_ARG, "non-finite tolerance");
  return PhaseObstacleRefitMidstep(ctx, inst, tolerance);
}

int xrt_obstacle_predicted_to_positions(xrt_ctx* ctx, int32_t inst) {
  int rc = ColliderInst(ctx, inst, true);
  if (rc != XRT_OK) return rc;
  ContactData& c = *ctx->contact[inst];
  XRT_CHECK_HIP(hipMemcpyAsync(c.o_x0.ptr, c.o_x.ptr, sizeof(float4) * c.obs_nv,
                               hipMemcpyDeviceToDevice, ctx->stream));
  return XRT_OK;
}

int xrt_vf_dcd_generate(xrt_ctx* ctx, int32_t inst) {
  int rc = ColliderInst(ctx, inst, true);
  return rc != XRT_OK ? rc : PhaseGenerate(ctx, inst, 0);
}

int xrt_vf_ccd_generate(xrt_ctx* ctx, int32_t inst) {
  int rc = ColliderInst(ctx, inst, true);
  return rc != XRT_OK ? rc : PhaseGenerate(ctx, inst, 1);
}

int xrt_ee_dcd_generate(xrt_ctx* ctx, int32_t inst) {
  int rc = ColliderInst(ctx, inst, true);
  return rc != XRT_OK ? rc : PhaseGenerate(ctx, inst, 2);
}

int xrt_vf_dcd_solve(xrt_ctx* ctx, int32_t inst) {
  int rc = ColliderInst(ctx, inst, true);
  if (rc != XRT_OK) return rc;
  ContactData& c = *ctx->contact[inst];
  const InstDesc& d = ctx->inst[inst].desc;
  XRT_CHECK_HIP(LaunchVfSolve(ctx->stream, ctx->x.ptr + d.vbeg, d.nv, c.obs_bvh, c.vf_face.ptr,
                              c.vf_q.ptr, 1e-3f, Friction(ctx, inst)));
  return XRT_OK;
}

int xrt_vf_ccd_solve(xrt_ctx* ctx, int32_t inst) {
  int rc = ColliderInst(ctx, inst, true);
  if (rc != XRT_OK) return rc;
  ContactData& c = *ctx->contact[inst];
  const InstDesc& d = ctx->inst[inst].desc;
  XRT_CHECK_HIP(LaunchVfSolve(ctx->stream, ctx->x.ptr + d.vbeg, d.nv, c.obs_bvh, c.rt_face.ptr,
                              c.rt_q.ptr, 0.0f, Friction(ctx, inst)));
  return XRT_OK;
}

int xrt_ee_dcd_solve(xrt_ctx* ctx, int32_t inst) {
  int rc = ColliderInst(ctx, inst, true);
  if (rc != XRT_OK) return rc;
  ContactData& c = *ctx->contact[inst];
  const HostInstance& h = ctx->inst[inst];
  XRT_CHECK_HIP(LaunchEeSolve(ctx->stream, ctx->x.ptr + h.desc.vbeg, c.c_edges.ptr, h.ne(),
                              reinterpret_cast<EeRec*>(c.ee.ptr), c.ee_none.ptr, c.dq.ptr,
                              c.cnt.ptr));
  return XRT_OK;
}

int xrt_repulsion_update_pairs(xrt_ctx* ctx, int32_t inst, float thickness) {
  int rc = ColliderInst(ctx, inst, false);
  if (rc != XRT_OK) return rc;
  XRT_REQUIRE(std::isfinite(thickness) && thickness >= 0.0f, XRT_ERR_INVALID_ARG,
              "thickness must be finite and >= 0");
  return PhaseUpdatePairs(ctx, inst, thickness);
}

int xrt_repulsion_solve_no_proximity(xrt_ctx* ctx, int32_t inst, float thickness,
                                     float relaxation_rate) {
  int rc = ColliderInst(ctx, inst, false);
  if (rc != XRT_OK) return rc;
  XRT_REQUIRE(std::isfinite(thickness) && std::isfinite(relaxation_rate), XRT_ERR_INVALID_ARG,
              "non-finite thickness or relaxation rate");
  return PhaseRepulsion(ctx, inst, thickness, relaxation_rate);
}

}  // extern "C"
