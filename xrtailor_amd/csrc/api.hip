// C ABI of the cloth-substep solver: context lifetime, scene setup, state I/O and the
// plain / batched substep drivers.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <string>

#include "context.h"
#include "dynamics.h"
#include "topology.h"

namespace xrt {
namespace {
thread_local std::string g_last_error;
}  // namespace

StepConsts MakeConsts(const xrt_params& p) {
  StepConsts k{};
  const float dt = p.frame_time / static_cast<float>(p.num_substeps);
  k.gx = p.gravity[0];
  k.gy = p.gravity[1];
  k.gz = p.gravity[2];
  k.dt = dt;
  k.inv_dt2 = 1.0f / (dt * dt);
  k.damp_factor = 1.0f - p.damping * dt;
  k.stretchiness = p.long_range_stretchiness;
  k.ground_y = p.ground_height;
  k.friction = p.ground_friction;
  k.relaxation = p.relaxation_factor;
  k.iterations = p.num_iterations;
  k.solve_bending = p.solve_bending;
  k.enable_ground = p.enable_ground;
  const bool consistent = p.bend_model == XRT_BEND_CONSISTENT;
  k.bend_dlo = consistent ? -1.0f : 0.0f;
  k.bend_th0 = consistent ? 3.14159265358979f : 0.0f;
  k.bend_th1 = consistent ? -1.0f : 1.0f;
  k.bend_sgn = consistent ? -1.0f : 1.0f;
  return k;
}

BendMode ModeOf(const StepConsts& k) { return BendMode{k.bend_dlo, k.bend_th0, k.bend_th1, k.bend_sgn}; }

// Sort constraints by (color, original index); returns local color offsets.
void ColorSort(const std::vector<int32_t>& color, int ncol, std::vector<int32_t>* order,
               std::vector<int32_t>* coloff) {
  const int n = static_cast<int>(color.size());
  coloff->assign(ncol + 1, 0);
  for (int i = 0; i < n; ++i) (*coloff)[color[i] + 1]++;
  for (int c = 0; c < ncol; ++c) (*coloff)[c + 1] += (*coloff)[c];
  std::vector<int32_t> fill(coloff->begin(), coloff->end() - 1);
  order->assign(n, 0);
  for (int i = 0; i < n; ++i) (*order)[fill[color[i]]++] = i;
}

namespace {
void DestroyGraph(xrt_ctx* c) {
  if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  c->graph_exec = nullptr;
  c->graph = nullptr;
}

int CheckInst(xrt_ctx* c, int32_t inst) {
  XRT_REQUIRE(c != nullptr, XRT_ERR_INVALID_ARG, "null context");
  XRT_REQUIRE(inst >= 0 && inst < static_cast<int32_t>(c->inst.size()), XRT_ERR_INVALID_ARG,
              "instance index out of range");
  return XRT_OK;
}

int CheckStarted(xrt_ctx* c, int32_t inst) {
  int rc = CheckInst(c, inst);
  if (rc != XRT_OK) return rc;
  XRT_REQUIRE(c->started, XRT_ERR_STATE, "xrt_start has not been called");
  return XRT_OK;
}


int UploadState(xrt_ctx* c, int32_t inst, const float* x0, const float* v) {
  xrt::HostInstance& h = c->inst[inst];
  const int nv = h.nv;
  if (x0) {
    std::vector<float4> buf(nv);
    for (int i = 0; i < nv; ++i)
      buf[i] = make_float4(x0[3 * i], x0[3 * i + 1], x0[3 * i + 2], h.inv_mass[i]);
    XRT_CHECK_HIP(hipMemcpyAsync(c->x0.ptr + h.desc.vbeg, buf.data(), sizeof(float4) * nv,
                                 hipMemcpyHostToDevice, c->stream));
    XRT_CHECK_HIP(hipMemcpyAsync(c->x.ptr + h.desc.vbeg, buf.data(), sizeof(float4) * nv,
                                 hipMemcpyHostToDevice, c->stream));
    XRT_CHECK_HIP(hipStreamSynchronize(c->stream));
  }
  if (v) {
    std::vector<float4> buf(nv);
    for (int i = 0; i < nv; ++i) buf[i] = make_float4(v[3 * i], v[3 * i + 1], v[3 * i + 2], 0.0f);
    XRT_CHECK_HIP(hipMemcpyAsync(c->v.ptr + h.desc.vbeg, buf.data(), sizeof(float4) * nv,
                                 hipMemcpyHostToDevice, c->stream));
    XRT_CHECK_HIP(hipStreamSynchronize(c->stream));
  }
  return XRT_OK;
}

int Download3(xrt_ctx* c, const float4* dev, int n, float* out) {
  std::vector<float4> buf(n);
  XRT_CHECK_HIP(hipMemcpyAsync(buf.data(), dev, sizeof(float4) * n, hipMemcpyDeviceToHost, c->stream));
  XRT_CHECK_HIP(hipStreamSynchronize(c->stream));
  for (int i = 0; i < n; ++i) {
    out[3 * i] = buf[i].x;
    out[3 * i + 1] = buf[i].y;
    out[3 * i + 2] = buf[i].z;
  }
  return XRT_OK;
}

BatchArgs MakeBatchArgs(xrt_ctx* c) {
  BatchArgs a{};
  a.inst = c->desc.ptr;
  a.order = c->order.ptr;
  a.x0 = c->x0.ptr;
  a.v = c->v.ptr;
  a.x = c->x.ptr;
  a.sc = c->sc.ptr;
  a.slam = c->slam.ptr;
  a.bidx = c->bidx.ptr;
  a.bpar = c->bpar.ptr;
  a.s_coloff = c->s_coloff.ptr;
  a.b_coloff = c->b_coloff.ptr;
  a.slot_vert = c->slot_vert.ptr;
  a.lra_ptr = c->lra_ptr.ptr;
  a.lra_slot = c->lra_slot.ptr;
  a.lra_rest = c->lra_rest.ptr;
  a.lds_verts = c->lds_verts;
  a.max_slots = c->max_slots;
  if (FemOn(c) && c->fem.ready) {
    const FemData& F = c->fem;
    a.fem = 1;
    a.fcol = F.fcol.ptr;
    a.f_coloff = F.f_coloff.ptr;
    a.fsc = F.sc.ptr;
    a.iso_k = F.iso_k.ptr;
    a.iso_c = F.iso_c.ptr;
    a.iso_lam = F.iso_lam.ptr;
    a.fk = FemConstsOf(c->params);
  }
  return a;
}

int SubstepPlain(xrt_ctx* c, int32_t inst) {
  const HostInstance& h = c->inst[inst];
  const InstDesc& d = h.desc;
  const StepConsts k = MakeConsts(c->params);
  if (c->contact[inst]) return SubstepPlainContact(c, inst, k, ModeOf(k));
  hipStream_t s = c->stream;
  float4* x0 = c->x0.ptr + d.vbeg;
  float4* v = c->v.ptr + d.vbeg;
  float4* x = c->x.ptr + d.vbeg;
  XRT_CHECK_HIP(LaunchPredict(s, x0, v, x, d.nv, k));
  int rc = ParticleSubstepBegin(c, inst);
  if (rc != XRT_OK) return rc;
  for (int it = 0; it < k.iterations; ++it) {
    if ((rc = ParticleIter(c, inst)) != XRT_OK) return rc;
    if (FemOn(c)) {
      if ((rc = FemIter(c, inst, it)) != XRT_OK) return rc;
    } else {
      for (int col = 0; col < h.s_ncol; ++col) {
        const int b = h.sbeg + h.s_coloff[col];
        XRT_CHECK_HIP(LaunchStretch(s, c->sc.ptr + b, c->slam.ptr + b, x,
                                    h.s_coloff[col + 1] - h.s_coloff[col], k.inv_dt2, it));
      }
    }
    if (k.solve_bending && !FemOn(c)) {
      for (int col = 0; col < h.b_ncol; ++col) {
        const int b = h.bbeg + h.b_coloff[col];
        XRT_CHECK_HIP(LaunchBend(s, c->bidx.ptr + b, c->bpar.ptr + b, x,
                                 h.b_coloff[col + 1] - h.b_coloff[col], k.inv_dt2, ModeOf(k)));
      }
    }
    if (d.nslots > 0)
      XRT_CHECK_HIP(LaunchLra(s, x, c->slot_vert.ptr + d.slot_beg, c->slot_pos.ptr, d.nslots,
                              c->lra_ptr.ptr + d.lra_vbeg, c->lra_slot.ptr, c->lra_rest.ptr, d.nv,
                              k.stretchiness));
    if (k.enable_ground) XRT_CHECK_HIP(LaunchGround(s, x, x0, d.nv, k.ground_y, k.friction));
  }
  XRT_CHECK_HIP(LaunchFinalize(s, x0, v, x, d.nv, k));
  return ParticleSubstepEnd(c, inst);
}

// Without contacts: one persistent launch (XPBD or FEM constraints).  With body or self
// contact, particle collision or bindings: the wide substep (each color, contact phase, hash
// step, binding pass and LBVH operation of every instance in one launch).
int SubstepBatched(xrt_ctx* c) {
  const StepConsts k = MakeConsts(c->params);
  XRT_REQUIRE(!FemOn(c) || c->fem.ready, XRT_ERR_STATE, "FEM constraints are not set up");
  // contacts, particles and bindings need the wide substep; without them it is a tuning
  // choice (every instance split over as many blocks as its colors need, one launch per
  // color, against one workgroup per instance in the persistent kernel)
  if (ContactsActive(c) || ParticlesOn(c) || BindingsOn(c) || c->tuning.path == 1)
    return WideSubstep(c, k, ModeOf(k), nullptr, -1);
  XRT_CHECK_HIP(LaunchBatchedSubstep(c->stream, MakeBatchArgs(c), k,
                                     static_cast<int>(c->inst.size()), c->tuning.persistent_block,
                                     c->batched_lds, c->tuning.persistent_unroll));
  return ParticleSubstepEnd(c, -1);
}

// The host-side frame schedule that decides which phases a batched frame launches (frame
// start, collision passes, particle hashes).  A frame of substeps brings it back to the same
// value, so a captured frame replays correctly exactly while the schedule equals the one at
// capture; single substeps (xrt_substep_batched / xrt_substep_plain) shift it.
std::vector<int> ScheduleKey(const xrt_ctx* c) {
  std::vector<int> key;
  for (const auto& cd : c->contact) {
    key.push_back(cd ? cd->frame_substep : -1);
    key.push_back(cd ? cd->pass_cnt : -1);
  }
  for (int s : c->particles.substep) key.push_back(s);
  return key;
}

int BuildGraph(xrt_ctx* c) {
  DestroyGraph(c);
  c->graph_key = ScheduleKey(c);
  XRT_CHECK_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  int rc = XRT_OK;
  for (int s = 0; s < c->params.num_substeps && rc == XRT_OK; ++s) rc = SubstepBatched(c);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(c->stream, &g);
  if (rc != XRT_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  XRT_CHECK_HIP(e);
  c->graph = g;
  XRT_CHECK_HIP(hipGraphInstantiate(&c->graph_exec, g, nullptr, nullptr, 0));
  return XRT_OK;
}

// Everything a run of batched substeps changes outside the iterate, saved so that the
// autotuner can time candidate launch shapes on the context's own workload and put the
// context back where it was: the state (x0, v), each contact instance's frame position,
// obstacle frame-start positions and contact records (solved until the next collision
// pass), the particle hash schedule and neighbor lists (used until the next hash).
struct SavedRecords {
  DevBuf<int32_t> vf_face, rt_face;
  DevBuf<float4> vf_q, rt_q;
  DevBuf<char> ee;
  DevBuf<uint8_t> ee_none;
};

struct SavedRun {
  DevBuf<float4> x0, v;
  std::vector<std::unique_ptr<DevBuf<float4>>> o_x0;
  std::vector<std::unique_ptr<SavedRecords>> rec;
  std::vector<int> frame_substep, pass_cnt;
  std::vector<char> moving;
  std::vector<int> particle_substep;
  DevBuf<uint32_t> nb;
};

template <class T>
hipError_t CopyBuf(DevBuf<T>& dst, const DevBuf<T>& src, hipStream_t s, bool alloc) {
  if (alloc) {
    const hipError_t e = dst.Alloc(src.n);
    if (e != hipSuccess) return e;
  }
  if (src.n == 0) return hipSuccess;
  return hipMemcpyAsync(dst.ptr, src.ptr, sizeof(T) * src.n, hipMemcpyDeviceToDevice, s);
}

int CopyRecords(SavedRecords& dst, const ContactData& cd, hipStream_t s) {
  XRT_CHECK_HIP(CopyBuf(dst.vf_face, cd.vf_face, s, true));
  XRT_CHECK_HIP(CopyBuf(dst.rt_face, cd.rt_face, s, true));
  XRT_CHECK_HIP(CopyBuf(dst.vf_q, cd.vf_q, s, true));
  XRT_CHECK_HIP(CopyBuf(dst.rt_q, cd.rt_q, s, true));
  XRT_CHECK_HIP(CopyBuf(dst.ee, cd.ee, s, true));
  XRT_CHECK_HIP(CopyBuf(dst.ee_none, cd.ee_none, s, true));
  return XRT_OK;
}

int RestoreRecords(ContactData& cd, const SavedRecords& src, hipStream_t s) {
  XRT_CHECK_HIP(CopyBuf(cd.vf_face, src.vf_face, s, false));
  XRT_CHECK_HIP(CopyBuf(cd.rt_face, src.rt_face, s, false));
  XRT_CHECK_HIP(CopyBuf(cd.vf_q, src.vf_q, s, false));
  XRT_CHECK_HIP(CopyBuf(cd.rt_q, src.rt_q, s, false));
  XRT_CHECK_HIP(CopyBuf(cd.ee, src.ee, s, false));
  XRT_CHECK_HIP(CopyBuf(cd.ee_none, src.ee_none, s, false));
  return XRT_OK;
}

int SaveRun(xrt_ctx* c, SavedRun& r) {
  const size_t n = c->x0.n;
  hipStream_t s = c->stream;
  XRT_CHECK_HIP(r.x0.Alloc(n));
  XRT_CHECK_HIP(r.v.Alloc(n));
  XRT_CHECK_HIP(hipMemcpyAsync(r.x0.ptr, c->x0.ptr, sizeof(float4) * n, hipMemcpyDeviceToDevice, s));
  XRT_CHECK_HIP(hipMemcpyAsync(r.v.ptr, c->v.ptr, sizeof(float4) * n, hipMemcpyDeviceToDevice, s));
  for (auto& cd : c->contact) {
    r.frame_substep.push_back(cd ? cd->frame_substep : 0);
    r.pass_cnt.push_back(cd ? cd->pass_cnt : 0);
    r.moving.push_back(cd ? cd->obstacle_moving : false);
    auto b = std::make_unique<DevBuf<float4>>();
    if (cd && cd->obs_nv > 0) {
      XRT_CHECK_HIP(b->Alloc(cd->obs_nv));
      XRT_CHECK_HIP(hipMemcpyAsync(b->ptr, cd->o_x0.ptr, sizeof(float4) * cd->obs_nv,
                                   hipMemcpyDeviceToDevice, s));
    }
    r.o_x0.push_back(std::move(b));
    auto rec = std::make_unique<SavedRecords>();
    if (cd) {
      const int rc = CopyRecords(*rec, *cd, s);
      if (rc != XRT_OK) return rc;
    }
    r.rec.push_back(std::move(rec));
  }
  r.particle_substep = c->particles.substep;
  if (c->particles.sized) XRT_CHECK_HIP(CopyBuf(r.nb, c->particles.nb, s, true));
  XRT_CHECK_HIP(hipStreamSynchronize(s));
  return XRT_OK;
}

int RestoreRun(xrt_ctx* c, const SavedRun& r) {
  const size_t n = c->x0.n;
  hipStream_t s = c->stream;
  XRT_CHECK_HIP(hipMemcpyAsync(c->x0.ptr, r.x0.ptr, sizeof(float4) * n, hipMemcpyDeviceToDevice, s));
  XRT_CHECK_HIP(hipMemcpyAsync(c->x.ptr, r.x0.ptr, sizeof(float4) * n, hipMemcpyDeviceToDevice, s));
  XRT_CHECK_HIP(hipMemcpyAsync(c->v.ptr, r.v.ptr, sizeof(float4) * n, hipMemcpyDeviceToDevice, s));
  for (size_t k = 0; k < c->contact.size(); ++k) {
    ContactData* cd = c->contact[k].get();
    if (!cd) continue;
    cd->frame_substep = r.frame_substep[k];
    cd->pass_cnt = r.pass_cnt[k];
    cd->obstacle_moving = r.moving[k] != 0;
    if (cd->obs_nv > 0)
      XRT_CHECK_HIP(hipMemcpyAsync(cd->o_x0.ptr, r.o_x0[k]->ptr, sizeof(float4) * cd->obs_nv,
                                   hipMemcpyDeviceToDevice, s));
    const int rc = RestoreRecords(*cd, *r.rec[k], s);
    if (rc != XRT_OK) return rc;
  }
  c->particles.substep = r.particle_substep;
  if (c->particles.sized && r.nb.n == c->particles.nb.n)
    XRT_CHECK_HIP(CopyBuf(c->particles.nb, r.nb, s, false));
  XRT_CHECK_HIP(hipStreamSynchronize(s));
  return XRT_OK;
}

// ms per batched substep with the context's current tuning, over n substeps (one untimed
// substep first), the run saved in r restored before and after
int TimeCandidate(xrt_ctx* c, const SavedRun& r, int n, float* ms) {
  int rc = RestoreRun(c, r);
  if (rc != XRT_OK) return rc;
  if ((rc = SubstepBatched(c)) != XRT_OK) return rc;  // warm-up (code load, caches)
  if ((rc = RestoreRun(c, r)) != XRT_OK) return rc;
  XRT_CHECK_HIP(hipEventRecord(c->ev0, c->stream));
  for (int i = 0; i < n; ++i)
    if ((rc = SubstepBatched(c)) != XRT_OK) return rc;
  XRT_CHECK_HIP(hipEventRecord(c->ev1, c->stream));
  XRT_CHECK_HIP(hipEventSynchronize(c->ev1));
  XRT_CHECK_HIP(hipEventElapsedTime(ms, c->ev0, c->ev1));
  *ms /= static_cast<float>(n);
  return RestoreRun(c, r);
}

int CheckTuning(const xrt_tuning& t) {
  XRT_REQUIRE(t.persistent_block >= 64 && t.persistent_block <= 1024 && t.persistent_block % 64 == 0,
              XRT_ERR_INVALID_ARG, "persistent_block must be a multiple of 64 in [64, 1024]");
  XRT_REQUIRE(t.persistent_unroll >= 0 && t.persistent_unroll <= 2, XRT_ERR_INVALID_ARG,
              "persistent_unroll must be 0, 1 or 2");
  XRT_REQUIRE(t.wide_per_thread == 1 || t.wide_per_thread == 2, XRT_ERR_INVALID_ARG,
              "wide_per_thread must be 1 or 2");
  XRT_REQUIRE(t.path == 0 || t.path == 1, XRT_ERR_INVALID_ARG, "path must be 0 or 1");
  return XRT_OK;
}

}  // namespace

void SetError(const std::string& msg) { g_last_error = msg; }

}  // namespace xrt

using namespace xrt;

extern "C" {

int xrt_abi_version(void) { return XRT_ABI_VERSION; }

const char* xrt_last_error(void) { return g_last_error.c_str(); }

void xrt_default_params(xrt_params* p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->gravity[0] = 0.0f;
  p->gravity[1] = -9.81f;
  p->gravity[2] = 0.0f;
  p->damping = 0.98f;
  p->relaxation_factor = 0.25f;
  p->long_range_stretchiness = 1.0f;
  p->bvh_tolerance = 0.003f;
  p->ground_height = -1.0f;
  p->ground_friction = 0.0f;
  p->pbd_thickness = 0.002f;
  p->repulsion_relaxation = 0.25f;
  p->frame_time = 1.0f / 60.0f;
  p->num_substeps = 10;
  p->num_iterations = 10;
  p->num_collision_passes = 10;
  p->solve_bending = 1;
  p->enable_ground = 1;
  p->enable_obstacle_contact = 0;
  p->enable_self_contact = 0;
  p->pair_update_interval = 10;
  p->bend_model = XRT_BEND_REFERENCE;
  p->coloring = XRT_COLOR_GREEDY;
  p->pair_capacity_per_face = 64;
  p->repulsion_vf_items_per_face = 64;
  p->repulsion_ee_items_per_face = 128;
  p->enable_particle_collision = 0;
  p->max_num_neighbors = 64;
  p->inter_leaved_hash = 3;
  p->particle_diameter = 1.3f;
  p->hash_cell_size = 1.5f;
  p->constraint_model = XRT_CONSTRAINTS_XPBD;
  p->fem_xx_stiffness = 1.0f;
  p->fem_yy_stiffness = 1.0f;
  p->fem_xy_stiffness = 1.0f;
  p->fem_xy_poisson_ratio = 0.3f;
  p->fem_yx_poisson_ratio = 0.3f;
  p->fem_bending_stiffness = 1.0f;
  p->contact_friction = 0.0f;
}

int xrt_create(xrt_ctx** out, int device) {
  XRT_REQUIRE(out != nullptr, XRT_ERR_INVALID_ARG, "null output pointer");
  *out = nullptr;
  int ndev = 0;
  XRT_CHECK_HIP(hipGetDeviceCount(&ndev));
  XRT_REQUIRE(device >= 0 && device < ndev, XRT_ERR_INVALID_ARG, "device index out of range");
  XRT_CHECK_HIP(hipSetDevice(device));
  auto* c = new xrt_ctx();
  c->device = device;
  xrt_default_params(&c->params);
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess ||
      hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
    SetError("device query / stream creation failed");
    delete c;
    return XRT_ERR_HIP;
  }
  c->cu_count = prop.multiProcessorCount;
  c->max_lds = prop.sharedMemPerBlock;
  int lds_optin = 0;
  if (hipDeviceGetAttribute(&lds_optin, hipDeviceAttributeSharedMemPerBlockOptin, device) ==
          hipSuccess &&
      static_cast<size_t>(lds_optin) > c->max_lds)
    c->max_lds = static_cast<size_t>(lds_optin);
  c->wave_size = prop.warpSize;
  c->l2_bytes = prop.l2CacheSize;
  *out = c;
  return XRT_OK;
}

int xrt_destroy(xrt_ctx* ctx) {
  if (!ctx) return XRT_OK;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  ctx->record.Release();
  DestroyGraph(ctx);
  if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
  if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return XRT_OK;
}

int xrt_set_params(xrt_ctx* ctx, const xrt_params* p) {
  XRT_REQUIRE(ctx && p, XRT_ERR_INVALID_ARG, "null argument");
  XRT_REQUIRE(p->num_substeps > 0 && p->num_iterations >= 0 && p->frame_time > 0.0f,
              XRT_ERR_INVALID_ARG, "num_substeps must be > 0, num_iterations >= 0, frame_time > 0");
  XRT_REQUIRE(p->num_collision_passes > 0, XRT_ERR_INVALID_ARG, "num_collision_passes must be > 0");
  XRT_REQUIRE(p->bend_model == XRT_BEND_REFERENCE || p->bend_model == XRT_BEND_CONSISTENT,
              XRT_ERR_INVALID_ARG, "unknown bend_model");
  XRT_REQUIRE(p->coloring == XRT_COLOR_GREEDY || p->coloring == XRT_COLOR_PALETTE,
              XRT_ERR_INVALID_ARG, "unknown coloring mode");
  XRT_REQUIRE(p->repulsion_vf_items_per_face >= 0 && p->repulsion_ee_items_per_face >= 0,
              XRT_ERR_INVALID_ARG, "repulsion item capacities must be >= 0");
  XRT_REQUIRE(std::isfinite(p->gravity[0]) && std::isfinite(p->gravity[1]) &&
                  std::isfinite(p->gravity[2]) && std::isfinite(p->damping),
              XRT_ERR_INVALID_ARG, "non-finite parameter");
  XRT_REQUIRE(p->max_num_neighbors >= 1 && p->max_num_neighbors <= 1024 && p->inter_leaved_hash >= 1,
              XRT_ERR_INVALID_ARG, "max_num_neighbors must be in [1, 1024], inter_leaved_hash >= 1");
  XRT_REQUIRE(p->particle_diameter > 0.0f && p->hash_cell_size > 0.0f &&
                  std::isfinite(p->particle_diameter) && std::isfinite(p->hash_cell_size),
              XRT_ERR_INVALID_ARG, "particle_diameter and hash_cell_size must be finite and > 0");
  XRT_REQUIRE(p->constraint_model == XRT_CONSTRAINTS_XPBD || p->constraint_model == XRT_CONSTRAINTS_FEM,
              XRT_ERR_INVALID_ARG, "unknown constraint_model");
  if (p->constraint_model == XRT_CONSTRAINTS_FEM) {
    const float fem[6] = {p->fem_xx_stiffness, p->fem_yy_stiffness, p->fem_xy_stiffness,
                          p->fem_xy_poisson_ratio, p->fem_yx_poisson_ratio, p->fem_bending_stiffness};
    for (float f : fem) XRT_REQUIRE(std::isfinite(f), XRT_ERR_INVALID_ARG, "non-finite FEM parameter");
    XRT_REQUIRE(p->fem_xx_stiffness >= 0.0f && p->fem_yy_stiffness >= 0.0f &&
                    p->fem_xy_stiffness >= 0.0f && p->fem_bending_stiffness > 0.0f,
                XRT_ERR_INVALID_ARG, "FEM stiffnesses must be >= 0, bending stiffness > 0");
    XRT_REQUIRE(p->fem_xy_poisson_ratio * p->fem_yx_poisson_ratio != 1.0f, XRT_ERR_INVALID_ARG,
                "fem_xy_poisson_ratio * fem_yx_poisson_ratio must differ from 1");
  }
  XRT_REQUIRE(p->pair_capacity_per_face >= 1, XRT_ERR_INVALID_ARG, "pair_capacity_per_face must be >= 1");
  XRT_REQUIRE(std::isfinite(p->contact_friction) && p->contact_friction >= 0.0f, XRT_ERR_INVALID_ARG,
              "contact_friction must be finite and >= 0");
  const bool lists = p->pair_capacity_per_face != ctx->params.pair_capacity_per_face ||
                     p->repulsion_vf_items_per_face != ctx->params.repulsion_vf_items_per_face ||
                     p->repulsion_ee_items_per_face != ctx->params.repulsion_ee_items_per_face;
  ctx->params = *p;
  DestroyGraph(ctx);
  if (!ctx->started) return XRT_OK;
  int rc = FemPrepare(ctx);
  if (rc != XRT_OK) return rc;
  if (lists && (rc = ContactResizeLists(ctx)) != XRT_OK) return rc;
  return ParticleConfigure(ctx);
}

int xrt_get_params(xrt_ctx* ctx, xrt_params* p) {
  XRT_REQUIRE(ctx && p, XRT_ERR_INVALID_ARG, "null argument");
  *p = ctx->params;
  return XRT_OK;
}

int xrt_sync(xrt_ctx* ctx) {
  XRT_REQUIRE(ctx, XRT_ERR_INVALID_ARG, "null context");
  return CheckPairOverflow(ctx);
}

int xrt_device_info(xrt_ctx* ctx, int64_t info[4]) {
  XRT_REQUIRE(ctx && info, XRT_ERR_INVALID_ARG, "null argument");
  info[0] = ctx->cu_count;
  info[1] = static_cast<int64_t>(ctx->max_lds);
  info[2] = ctx->wave_size;
  info[3] = ctx->l2_bytes;
  return XRT_OK;
}

int xrt_build_edges(const int32_t* tris, int32_t nt, int32_t* edges_out, int32_t max_edges,
                    int32_t* ne_out) {
  XRT_REQUIRE(tris && nt >= 0 && ne_out, XRT_ERR_INVALID_ARG, "null argument");
  std::vector<int32_t> e = BuildEdges(tris, nt);
  *ne_out = static_cast<int32_t>(e.size() / 2);
  XRT_REQUIRE(*ne_out <= max_edges, XRT_ERR_CAPACITY, "edge buffer too small");
  if (edges_out) std::copy(e.begin(), e.end(), edges_out);
  return XRT_OK;
}

int xrt_build_bends(const int32_t* tris, int32_t nt, const int32_t* edges, int32_t ne,
                    int32_t* bends_out, int32_t max_bends, int32_t* nb_out) {
  XRT_REQUIRE(tris && edges && nt >= 0 && ne >= 0 && nb_out, XRT_ERR_INVALID_ARG, "null argument");
  std::vector<int32_t> ev(edges, edges + 2 * ne);
  std::vector<int32_t> b = BuildBends(tris, nt, ev);
  *nb_out = static_cast<int32_t>(b.size() / 4);
  XRT_REQUIRE(*nb_out <= max_bends, XRT_ERR_CAPACITY, "bend buffer too small");
  if (bends_out) std::copy(b.begin(), b.end(), bends_out);
  return XRT_OK;
}

int xrt_color_constraints(const int32_t* cons, int32_t arity, int32_t n, int32_t nv,
                          uint64_t seed, int32_t mode, int32_t* colors_out,
                          int32_t* ncolors_out) {
  XRT_REQUIRE(arity > 0 && n >= 0 && nv >= 0 && ncolors_out && (n == 0 || (cons && colors_out)),
              XRT_ERR_INVALID_ARG, "bad coloring arguments");
  XRT_REQUIRE(mode == XRT_COLOR_GREEDY || mode == XRT_COLOR_PALETTE, XRT_ERR_INVALID_ARG,
              "unknown coloring mode");
  for (int i = 0; i < n * arity; ++i)
    XRT_REQUIRE(cons[i] >= 0 && cons[i] < nv, XRT_ERR_INVALID_ARG, "vertex id out of range");
  int ncol = 0;
  std::vector<int32_t> col = ColorConstraints(cons, arity, n, nv, seed, mode, &ncol);
  std::copy(col.begin(), col.end(), colors_out);
  *ncolors_out = ncol;
  return XRT_OK;
}

int xrt_add_cloth(xrt_ctx* ctx, const float* pos, int32_t nv, const int32_t* tris, int32_t nt,
                  float stretch_compliance, float bend_compliance, int32_t* instance_out) {
  XRT_REQUIRE(ctx != nullptr, XRT_ERR_INVALID_ARG, "null context");
  XRT_REQUIRE(!ctx->started, XRT_ERR_STATE, "cannot add cloth after xrt_start");
  XRT_REQUIRE(nv > 0 && nt > 0 && pos && tris, XRT_ERR_INVALID_ARG, "empty or null mesh");
  XRT_REQUIRE(stretch_compliance >= 0.0f && bend_compliance >= 0.0f, XRT_ERR_INVALID_ARG,
              "compliance must be >= 0");
  for (int i = 0; i < 3 * nt; ++i)
    XRT_REQUIRE(tris[i] >= 0 && tris[i] < nv, XRT_ERR_INVALID_ARG, "triangle index out of range");
  for (int f = 0; f < nt; ++f) {
    const int a = tris[3 * f], b = tris[3 * f + 1], cc = tris[3 * f + 2];
    XRT_REQUIRE(a != b && b != cc && a != cc, XRT_ERR_INVALID_ARG, "degenerate triangle");
  }
  for (int i = 0; i < 3 * nv; ++i)
    XRT_REQUIRE(std::isfinite(pos[i]), XRT_ERR_INVALID_ARG, "non-finite vertex position");
  HostInstance h;
  h.nv = nv;
  h.nt = nt;
  h.rest_pos.assign(pos, pos + 3 * nv);
  h.tris.assign(tris, tris + 3 * nt);
  h.inv_mass.assign(nv, 1.0f);  // free nodes (InitializeNodeGeometriesLocal, PhysicsMeshHelper.cu:139)
  h.edges = BuildEdges(tris, nt);
  h.rest_len = RestLengths(pos, h.edges);
  h.s_comp.assign(h.ne(), stretch_compliance);
  h.bends = BuildBends(tris, nt, h.edges);
  h.theta = RestAngles(pos, h.bends);
  h.b_comp.assign(h.nb(), bend_compliance);
  ctx->inst.push_back(std::move(h));
  if (instance_out) *instance_out = static_cast<int32_t>(ctx->inst.size() - 1);
  return XRT_OK;
}

int xrt_pin_vertices(xrt_ctx* ctx, int32_t inst, const int32_t* idx, int32_t n) {
  int rc = CheckInst(ctx, inst);
  if (rc != XRT_OK) return rc;
  XRT_REQUIRE(!ctx->started, XRT_ERR_STATE, "pin vertices before xrt_start");
  XRT_REQUIRE(n >= 0 && (n == 0 || idx), XRT_ERR_INVALID_ARG, "null index array");
  HostInstance& h = ctx->inst[inst];
  for (int i = 0; i < n; ++i)
    XRT_REQUIRE(idx[i] >= 0 && idx[i] < h.nv, XRT_ERR_INVALID_ARG, "pinned index out of range");
  for (int i = 0; i < n; ++i) h.inv_mass[idx[i]] = 0.0f;
  return XRT_OK;
}

int xrt_set_stretch_compliance(xrt_ctx* ctx, int32_t inst, const float* compliance) {
  int rc = CheckInst(ctx, inst);
  if (rc != XRT_OK) return rc;
  XRT_REQUIRE(compliance, XRT_ERR_INVALID_ARG, "null compliance");
  HostInstance& h = ctx->inst[inst];
  for (int e = 0; e < h.ne(); ++e)
    XRT_REQUIRE(compliance[e] >= 0.0f && std::isfinite(compliance[e]), XRT_ERR_INVALID_ARG,
                "compliance must be finite and >= 0");
  h.s_comp.assign(compliance, compliance + h.ne());
  if (ctx->started) {
    std::vector<StretchC> sc(h.ne());
    for (int p = 0; p < h.ne(); ++p) {
      const int e = h.s_order[p];
      sc[p] = {h.edges[2 * e], h.edges[2 * e + 1], h.rest_len[e], h.s_comp[e]};
    }
    XRT_CHECK_HIP(hipMemcpyAsync(ctx->sc.ptr + h.sbeg, sc.data(), sizeof(StretchC) * sc.size(),
                                 hipMemcpyHostToDevice, ctx->stream));
    XRT_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  }
  return XRT_OK;
}

int xrt_add_lra_slots(xrt_ctx* ctx, int32_t inst, const int32_t* slots, int32_t n) {
  int rc = CheckInst(ctx, inst);
  if (rc != XRT_OK) return rc;
  XRT_REQUIRE(!ctx->started, XRT_ERR_STATE, "add LRA slots before xrt_start");
  XRT_REQUIRE(n >= 0 && (n == 0 || slots), XRT_ERR_INVALID_ARG, "null slot array");
  HostInstance& h = ctx->inst[inst];
  for (int i = 0; i < n; ++i)
    XRT_REQUIRE(slots[i] >= 0 && slots[i] < h.nv, XRT_ERR_INVALID_ARG, "slot index out of range");
  for (int i = 0; i < n; ++i) {
    h.slots.push_back(slots[i]);
    h.slot_dist.push_back(Geodesic(h.rest_pos.data(), h.nv, h.edges, slots[i]));
  }
  return XRT_OK;
}

int xrt_start(xrt_ctx* ctx, uint64_t seed) {
  XRT_REQUIRE(ctx != nullptr, XRT_ERR_INVALID_ARG, "null context");
  XRT_REQUIRE(!ctx->started, XRT_ERR_STATE, "xrt_start called twice");
  XRT_REQUIRE(!ctx->inst.empty(), XRT_ERR_STATE, "no cloth instance registered");
  XRT_CHECK_HIP(hipSetDevice(ctx->device));
  int vtot = 0, stot = 0, btot = 0, scol = 0, bcol = 0, slot_tot = 0, lra_v = 0, ktot = 0;
  int max_slots = 0;
  for (size_t ii = 0; ii < ctx->inst.size(); ++ii) {
    HostInstance& h = ctx->inst[ii];
    const uint64_t iseed = seed + ii;
    const int mode = ctx->params.coloring;
    h.s_color = ColorConstraints(h.edges.data(), 2, h.ne(), h.nv, iseed, mode, &h.s_ncol);
    h.b_color = ColorConstraints(h.bends.data(), 4, h.nb(), h.nv, iseed, mode, &h.b_ncol);
    ColorSort(h.s_color, h.s_ncol, &h.s_order, &h.s_coloff);
    ColorSort(h.b_color, h.b_ncol, &h.b_order, &h.b_coloff);
    // per-target LRA CSR, slots in registration order; unreachable targets are dropped
    const int ns = static_cast<int>(h.slots.size());
    h.lra_ptr.assign(h.nv + 1, 0);
    h.lra_slot.clear();
    h.lra_rest.clear();
    for (int t = 0; t < h.nv; ++t) {
      for (int j = 0; j < ns; ++j) {
        const double dist = h.slot_dist[j][t];
        if (!std::isfinite(dist)) continue;
        h.lra_slot.push_back(j);
        h.lra_rest.push_back(static_cast<float>(dist));
      }
      h.lra_ptr[t + 1] = static_cast<int32_t>(h.lra_slot.size());
    }
    InstDesc& d = h.desc;
    d.vbeg = vtot;
    d.nv = h.nv;
    d.s_cbeg = scol;
    d.s_ncol = h.s_ncol;
    d.b_cbeg = bcol;
    d.b_ncol = h.b_ncol;
    d.slot_beg = slot_tot;
    d.nslots = ns;
    d.lra_vbeg = lra_v;
    h.sbeg = stot;
    h.bbeg = btot;
    h.kbeg = ktot;
    vtot += h.nv;
    stot += h.ne();
    btot += h.nb();
    scol += h.s_ncol + 1;
    bcol += h.b_ncol + 1;
    slot_tot += ns;
    lra_v += h.nv + 1;
    ktot += static_cast<int>(h.lra_slot.size());
    max_slots = std::max(max_slots, ns);
  }
  std::vector<float4> hx0(vtot), hv(vtot);
  std::vector<StretchC> hsc(stot);
  std::vector<int4> hbidx(btot);
  std::vector<float2> hbpar(btot);
  std::vector<int32_t> hsco(scol), hbco(bcol), hslot(slot_tot), hlptr(lra_v), hlslot(ktot);
  std::vector<float> hlrest(ktot);
  std::vector<InstDesc> hdesc(ctx->inst.size());
  for (size_t ii = 0; ii < ctx->inst.size(); ++ii) {
    const HostInstance& h = ctx->inst[ii];
    const InstDesc& d = h.desc;
    hdesc[ii] = d;
    for (int i = 0; i < h.nv; ++i) {
      const float* p = h.x0_init.empty() ? &h.rest_pos[3 * i] : &h.x0_init[3 * i];
      hx0[d.vbeg + i] = make_float4(p[0], p[1], p[2], h.inv_mass[i]);
      if (h.v_init.empty())
        hv[d.vbeg + i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      else
        hv[d.vbeg + i] = make_float4(h.v_init[3 * i], h.v_init[3 * i + 1], h.v_init[3 * i + 2], 0.0f);
    }
    for (int p = 0; p < h.ne(); ++p) {
      const int e = h.s_order[p];
      hsc[h.sbeg + p] = {h.edges[2 * e], h.edges[2 * e + 1], h.rest_len[e], h.s_comp[e]};
    }
    for (int p = 0; p < h.nb(); ++p) {
      const int b = h.b_order[p];
      hbidx[h.bbeg + p] = make_int4(h.bends[4 * b], h.bends[4 * b + 1], h.bends[4 * b + 2],
                                    h.bends[4 * b + 3]);
      hbpar[h.bbeg + p] = make_float2(h.theta[b], h.b_comp[b]);
    }
    for (int cidx = 0; cidx <= h.s_ncol; ++cidx) hsco[d.s_cbeg + cidx] = h.sbeg + h.s_coloff[cidx];
    for (int cidx = 0; cidx <= h.b_ncol; ++cidx) hbco[d.b_cbeg + cidx] = h.bbeg + h.b_coloff[cidx];
    for (int j = 0; j < d.nslots; ++j) hslot[d.slot_beg + j] = h.slots[j];
    for (int t = 0; t <= h.nv; ++t) hlptr[d.lra_vbeg + t] = h.kbeg + h.lra_ptr[t];
    for (size_t q = 0; q < h.lra_slot.size(); ++q) {
      hlslot[h.kbeg + q] = h.lra_slot[q];
      hlrest[h.kbeg + q] = h.lra_rest[q];
    }
  }
  // workgroup order: largest instance first (longest-processing-time first)
  std::vector<int32_t> order(ctx->inst.size());
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return ctx->inst[a].nv > ctx->inst[b].nv; });
  hipStream_t s = ctx->stream;
  XRT_CHECK_HIP(ctx->x0.Upload(hx0, s));
  XRT_CHECK_HIP(ctx->x.Upload(hx0, s));
  XRT_CHECK_HIP(ctx->v.Upload(hv, s));
  XRT_CHECK_HIP(ctx->sc.Upload(hsc, s));
  XRT_CHECK_HIP(ctx->slam.Alloc(std::max(stot, 1)));
  XRT_CHECK_HIP(ctx->bidx.Upload(hbidx, s));
  XRT_CHECK_HIP(ctx->bpar.Upload(hbpar, s));
  XRT_CHECK_HIP(ctx->s_coloff.Upload(hsco, s));
  XRT_CHECK_HIP(ctx->b_coloff.Upload(hbco, s));
  XRT_CHECK_HIP(ctx->slot_vert.Upload(hslot, s));
  XRT_CHECK_HIP(ctx->lra_ptr.Upload(hlptr, s));
  XRT_CHECK_HIP(ctx->lra_slot.Upload(hlslot, s));
  XRT_CHECK_HIP(ctx->lra_rest.Upload(hlrest, s));
  XRT_CHECK_HIP(ctx->slot_pos.Alloc(std::max(max_slots, 1)));
  ctx->total_slots = slot_tot;
  XRT_CHECK_HIP(ctx->slot_pos_all.Alloc(std::max(slot_tot, 1)));
  XRT_CHECK_HIP(ctx->desc.Upload(hdesc, s));
  XRT_CHECK_HIP(ctx->order.Upload(order, s));
  XRT_CHECK_HIP(hipMemsetAsync(ctx->slam.ptr, 0, sizeof(float) * std::max(stot, 1), s));
  XRT_CHECK_HIP(hipStreamSynchronize(s));
  // batched launch: the iterate x of an instance is staged in LDS when it fits
  ctx->max_slots = max_slots;
  const size_t lds_cap = std::min<size_t>(ctx->max_lds, 160 * 1024);
  const size_t slot_bytes = sizeof(float4) * static_cast<size_t>(max_slots);
  int lds_verts = 0;
  if (lds_cap > slot_bytes) lds_verts = static_cast<int>((lds_cap - slot_bytes) / sizeof(float4));
  int max_nv = 0;
  for (const auto& h : ctx->inst) max_nv = std::max(max_nv, h.nv);
  ctx->max_nv = max_nv;
  lds_verts = std::min(lds_verts, max_nv);
  ctx->lds_verts = lds_verts;
  ctx->batched_lds = sizeof(float4) * (static_cast<size_t>(lds_verts) + max_slots);
  if (ctx->batched_lds > 64 * 1024) XRT_CHECK_HIP(ConfigureBatchedKernel(ctx->batched_lds));
  // wide launches (one color of every instance): grid sized by the largest color
  WideConfig& w = ctx->wide;
  w.s_max.clear();
  w.b_max.clear();
  w.n_total = vtot;
  w.e_total = stot;
  w.b_total = btot;
  for (const auto& h : ctx->inst) {
    for (int col = 0; col < h.s_ncol; ++col) {
      if (col >= static_cast<int>(w.s_max.size())) w.s_max.push_back(0);
      w.s_max[col] = std::max(w.s_max[col], h.s_coloff[col + 1] - h.s_coloff[col]);
    }
    for (int col = 0; col < h.b_ncol; ++col) {
      if (col >= static_cast<int>(w.b_max.size())) w.b_max.push_back(0);
      w.b_max[col] = std::max(w.b_max[col], h.b_coloff[col + 1] - h.b_coloff[col]);
    }
  }
  {  // per-color block maps of the wide launches (only blocks with elements are launched)
    const int ni = static_cast<int>(ctx->inst.size());
    std::vector<int32_t> maps;
    auto build = [&](bool stretch, std::vector<size_t>& seg, std::vector<int>& tot) {
      const size_t ncol = stretch ? w.s_max.size() : w.b_max.size();
      seg.assign(ncol, 0);
      tot.assign(ncol, 0);
      for (size_t col = 0; col < ncol; ++col) {
        std::vector<int32_t> o(ni + 1, 0);
        for (int k = 0; k < ni; ++k) {
          const HostInstance& h = ctx->inst[k];
          const int nc = stretch ? h.s_ncol : h.b_ncol;
          const std::vector<int32_t>& co = stretch ? h.s_coloff : h.b_coloff;
          const int cnt = static_cast<int>(col) < nc ? co[col + 1] - co[col] : 0;
          o[k + 1] = o[k] + (cnt + 255) / 256;
        }
        seg[col] = maps.size();
        tot[col] = o[ni];
        maps.insert(maps.end(), o.begin(), o.end());
        for (int k = 0; k < ni; ++k) maps.insert(maps.end(), static_cast<size_t>(o[k + 1] - o[k]), k);
      }
    };
    build(true, w.s_map, w.s_blocks);
    build(false, w.b_map, w.b_blocks);
    XRT_CHECK_HIP(w.maps.Upload(maps, ctx->stream));
    XRT_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  }
  int rc = ContactStart(ctx);
  if (rc != XRT_OK) return rc;
  if ((rc = ParticleStart(ctx)) != XRT_OK) return rc;
  if ((rc = BindingStart(ctx)) != XRT_OK) return rc;
  ctx->fem.seed = seed;
  if ((rc = FemPrepare(ctx)) != XRT_OK) return rc;
  ctx->started = true;
  return XRT_OK;
}

int xrt_num_instances(xrt_ctx* ctx, int32_t* n) {
  XRT_REQUIRE(ctx && n, XRT_ERR_INVALID_ARG, "null argument");
  *n = static_cast<int32_t>(ctx->inst.size());
  return XRT_OK;
}

int xrt_get_counts(xrt_ctx* ctx, int32_t inst, int32_t counts[8]) {
  int rc = CheckInst(ctx, inst);
  if (rc != XRT_OK) return rc;
  XRT_REQUIRE(counts, XRT_ERR_INVALID_ARG, "null counts");
  const HostInstance& h = ctx->inst[inst];
  counts[0] = h.nv;
  counts[1] = h.nt;
  counts[2] = h.ne();
  counts[3] = h.nb();
  counts[4] = h.s_ncol;
  counts[5] = h.b_ncol;
  counts[6] = static_cast<int32_t>(h.slots.size());
  counts[7] = static_cast<int32_t>(h.lra_slot.size());
  return XRT_OK;
}

int xrt_get_edges(xrt_ctx* ctx, int32_t inst, int32_t* edges, float* rest, int32_t* colors) {
  int rc = CheckInst(ctx, inst);
  if (rc != XRT_OK) return rc;
  const HostInstance& h = ctx->inst[inst];
  if (edges) std::copy(h.edges.begin(), h.edges.end(), edges);
  if (rest) std::copy(h.rest_len.begin(), h.rest_len.end(), rest);
  if (colors) {
    XRT_REQUIRE(ctx->started, XRT_ERR_STATE, "colors exist after xrt_start");
    std::copy(h.s_color.begin(), h.s_color.end(), colors);
  }
  return XRT_OK;
}

int xrt_get_bends(xrt_ctx* ctx, int32_t inst, int32_t* bends, float* theta, int32_t* colors) {
  int rc = CheckInst(ctx, inst);
  if (rc != XRT_OK) return rc;
  const HostInstance& h = ctx->inst[inst];
  if (bends) std::copy(h.bends.begin(), h.bends.end(), bends);
  if (theta) std::copy(h.theta.begin(), h.theta.end(), theta);
  if (colors) {
    XRT_REQUIRE(ctx->started, XRT_ERR_STATE, "colors exist after xrt_start");
    std::copy(h.b_color.begin(), h.b_color.end(), colors);
  }
  return XRT_OK;
}

int xrt_get_lra(xrt_ctx* ctx, int32_t inst, int32_t* slot_vert, int32_t* ptr, int32_t* slot,
                float* rest) {
  int rc = CheckStarted(ctx, inst);
  if (rc != XRT_OK) return rc;
  const HostInstance& h = ctx->inst[inst];
  if (slot_vert) std::copy(h.slots.begin(), h.slots.end(), slot_vert);
  if (ptr) std::copy(h.lra_ptr.begin(), h.lra_ptr.end(), ptr);
  if (slot) std::copy(h.lra_slot.begin(), h.lra_slot.end(), slot);
  if (rest) std::copy(h.lra_rest.begin(), h.lra_rest.end(), rest);
  return XRT_OK;
}

int xrt_get_inv_mass(xrt_ctx* ctx, int32_t inst, float* w) {
  int rc = CheckInst(ctx, inst);
  if (rc != XRT_OK) return rc;
  XRT_REQUIRE(w, XRT_ERR_INVALID_ARG, "null output");
  std::copy(ctx->inst[inst].inv_mass.begin(), ctx->inst[inst].inv_mass.end(), w);
  return XRT_OK;
}

int xrt_set_state(xrt_ctx* ctx, int32_t inst, const float* x0, const float* v) {
  int rc = CheckInst(ctx, inst);
  if (rc != XRT_OK) return rc;
  HostInstance& h = ctx->inst[inst];
  if (x0)
    for (int i = 0; i < 3 * h.nv; ++i)
      XRT_REQUIRE(std::isfinite(x0[i]), XRT_ERR_NOT_FINITE, "non-finite position");
  if (v)
    for (int i = 0; i < 3 * h.nv; ++i)
      XRT_REQUIRE(std::isfinite(v[i]), XRT_ERR_NOT_FINITE, "non-finite velocity");
  if (!ctx->started) {
    if (x0) h.x0_init.assign(x0, x0 + 3 * h.nv);
    if (v) h.v_init.assign(v, v + 3 * h.nv);
    return XRT_OK;
  }
  rc = UploadState(ctx, inst, x0, v);
  if (rc != XRT_OK || !x0) return rc;
  if ((rc = ContactStateReset(ctx, inst)) != XRT_OK) return rc;
  return ParticleStateReset(ctx, inst);
}

int xrt_get_state(xrt_ctx* ctx, int32_t inst, float* x0, float* v) {
  int rc = CheckStarted(ctx, inst);
  if (rc != XRT_OK) return rc;
  const InstDesc& d = ctx->inst[inst].desc;
  if (x0 && (rc = Download3(ctx, ctx->x0.ptr + d.vbeg, d.nv, x0)) != XRT_OK) return rc;
  if (v && (rc = Download3(ctx, ctx->v.ptr + d.vbeg, d.nv, v)) != XRT_OK) return rc;
  return XRT_OK;
}

int xrt_get_predicted(xrt_ctx* ctx, int32_t inst, float* x) {
  int rc = CheckStarted(ctx, inst);
  if (rc != XRT_OK) return rc;
  XRT_REQUIRE(x, XRT_ERR_INVALID_ARG, "null output");
  const InstDesc& d = ctx->inst[inst].desc;
  return Download3(ctx, ctx->x.ptr + d.vbeg, d.nv, x);
}

int xrt_predict_positions(xrt_ctx* ctx, int32_t inst, float dt) {
  int rc = CheckStarted(ctx, inst);
  if (rc != XRT_OK) return rc;
  XRT_REQUIRE(dt > 0.0f, XRT_ERR_INVALID_ARG, "dt must be > 0");
  StepConsts k = MakeConsts(ctx->params);
  k.dt = dt;
  const InstDesc& d = ctx->inst[inst].desc;
  XRT_CHECK_HIP(LaunchPredict(ctx->stream, ctx->x0.ptr + d.vbeg, ctx->v.ptr + d.vbeg,
                              ctx->x.ptr + d.vbeg, d.nv, k));
  return XRT_OK;
}

int xrt_stretch_solve(xrt_ctx* ctx, int32_t inst, int32_t color, float dt, int32_t iter) {
  int rc = CheckStarted(ctx, inst);
  if (rc != XRT_OK) return rc;
  const HostInstance& h = ctx->inst[inst];
  XRT_REQUIRE(color >= 0 && color < h.s_ncol, XRT_ERR_INVALID_ARG, "color out of range");
  XRT_REQUIRE(dt > 0.0f, XRT_ERR_INVALID_ARG, "dt must be > 0");
  const int b = h.sbeg + h.s_coloff[color];
  XRT_CHECK_HIP(LaunchStretch(ctx->stream, ctx->sc.ptr + b, ctx->slam.ptr + b,
                              ctx->x.ptr + h.desc.vbeg, h.s_coloff[color + 1] - h.s_coloff[color],
                              1.0f / (dt * dt), iter));
  return XRT_OK;
}

int xrt_bend_solve(xrt_ctx* ctx, int32_t inst, int32_t color, float dt, int32_t iter) {
  (void)iter;
  int rc = CheckStarted(ctx, inst);
  if (rc != XRT_OK) return rc;
  const HostInstance& h = ctx->inst[inst];
  XRT_REQUIRE(color >= 0 && color < h.b_ncol, XRT_ERR_INVALID_ARG, "color out of range");
  XRT_REQUIRE(dt > 0.0f, XRT_ERR_INVALID_ARG, "dt must be > 0");
  const int b = h.bbeg + h.b_coloff[color];
  XRT_CHECK_HIP(LaunchBend(ctx->stream, ctx->bidx.ptr + b, ctx->bpar.ptr + b,
                           ctx->x.ptr + h.desc.vbeg, h.b_coloff[color + 1] - h.b_coloff[color],
                           1.0f / (dt * dt), ModeOf(MakeConsts(ctx->params))));
  return XRT_OK;
}

int xrt_solve_lra(xrt_ctx* ctx, int32_t inst) {
  int rc = CheckStarted(ctx, inst);
  if (rc != XRT_OK) return rc;
  const InstDesc& d = ctx->inst[inst].desc;
  XRT_CHECK_HIP(LaunchLra(ctx->stream, ctx->x.ptr + d.vbeg, ctx->slot_vert.ptr + d.slot_beg,
                          ctx->slot_pos.ptr, d.nslots, ctx->lra_ptr.ptr + d.lra_vbeg,
                          ctx->lra_slot.ptr, ctx->lra_rest.ptr, d.nv,
                          ctx->params.long_range_stretchiness));
  return XRT_OK;
}

int xrt_solve_ground(xrt_ctx* ctx, int32_t inst) {
  int rc = CheckStarted(ctx, inst);
  if (rc != XRT_OK) return rc;
  const InstDesc& d = ctx->inst[inst].desc;
  XRT_CHECK_HIP(LaunchGround(ctx->stream, ctx->x.ptr + d.vbeg, ctx->x0.ptr + d.vbeg, d.nv,
                             ctx->params.ground_height, ctx->params.ground_friction));
  return XRT_OK;
}

int xrt_apply_deltas(xrt_ctx* ctx, int32_t inst) {
  int rc = CheckStarted(ctx, inst);
  if (rc != XRT_OK) return rc;
  return ApplyInstanceDeltas(ctx, inst);
}

int xrt_finalize(xrt_ctx* ctx, int32_t inst, float dt) {
  int rc = CheckStarted(ctx, inst);
  if (rc != XRT_OK) return rc;
  XRT_REQUIRE(dt > 0.0f, XRT_ERR_INVALID_ARG, "dt must be > 0");
  StepConsts k = MakeConsts(ctx->params);
  k.dt = dt;
  k.damp_factor = 1.0f - ctx->params.damping * dt;
  const InstDesc& d = ctx->inst[inst].desc;
  XRT_CHECK_HIP(LaunchFinalize(ctx->stream, ctx->x0.ptr + d.vbeg, ctx->v.ptr + d.vbeg,
                               ctx->x.ptr + d.vbeg, d.nv, k));
  return XRT_OK;
}

int xrt_substep_plain(xrt_ctx* ctx, int32_t inst) {
  int rc = CheckStarted(ctx, inst);
  if (rc != XRT_OK) return rc;
  return SubstepPlain(ctx, inst);
}

int xrt_frames_plain(xrt_ctx* ctx, int32_t nframes) {
  XRT_REQUIRE(ctx && ctx->started, XRT_ERR_STATE, "xrt_start has not been called");
  XRT_REQUIRE(nframes >= 0, XRT_ERR_INVALID_ARG, "nframes must be >= 0");
  for (int f = 0; f < nframes; ++f)
    for (size_t i = 0; i < ctx->inst.size(); ++i)
      for (int s = 0; s < ctx->params.num_substeps; ++s) {
        int rc = SubstepPlain(ctx, static_cast<int32_t>(i));
        if (rc != XRT_OK) return rc;
      }
  return XRT_OK;
}

int xrt_substep_batched(xrt_ctx* ctx) {
  XRT_REQUIRE(ctx && ctx->started, XRT_ERR_STATE, "xrt_start has not been called");
  return SubstepBatched(ctx);
}

int xrt_frames_batched(xrt_ctx* ctx, int32_t nframes) {
  XRT_REQUIRE(ctx && ctx->started, XRT_ERR_STATE, "xrt_start has not been called");
  XRT_REQUIRE(nframes >= 0, XRT_ERR_INVALID_ARG, "nframes must be >= 0");
  if (!ctx->graph_exec || ScheduleKey(ctx) != ctx->graph_key) {
    int rc = BuildGraph(ctx);
    if (rc != XRT_OK) return rc;
  }
  for (int f = 0; f < nframes; ++f) XRT_CHECK_HIP(hipGraphLaunch(ctx->graph_exec, ctx->stream));
  return XRT_OK;
}

int xrt_get_tuning(xrt_ctx* ctx, xrt_tuning* out) {
  XRT_REQUIRE(ctx && out, XRT_ERR_INVALID_ARG, "null argument");
  *out = ctx->tuning;
  return XRT_OK;
}

int xrt_set_tuning(xrt_ctx* ctx, const xrt_tuning* t) {
  XRT_REQUIRE(ctx && t, XRT_ERR_INVALID_ARG, "null argument");
  const int rc = CheckTuning(*t);
  if (rc != XRT_OK) return rc;
  ctx->tuning.persistent_block = t->persistent_block;
  ctx->tuning.persistent_unroll = t->persistent_unroll;
  ctx->tuning.wide_per_thread = t->wide_per_thread;
  ctx->tuning.path = t->path;
  DestroyGraph(ctx);  // the next batched frame captures the new shapes
  return XRT_OK;
}

int xrt_autotune(xrt_ctx* ctx, int32_t substeps, xrt_tuning* out) {
  XRT_REQUIRE(ctx && ctx->started, XRT_ERR_STATE, "xrt_start has not been called");
  XRT_REQUIRE(substeps > 0, XRT_ERR_INVALID_ARG, "substeps must be > 0");
  XRT_REQUIRE(!FemOn(ctx) || ctx->fem.ready, XRT_ERR_STATE, "FEM constraints are not set up");
  int rc = CheckPairOverflow(ctx);  // also waits for queued work
  if (rc != XRT_OK) return rc;
  SavedRun saved;
  if ((rc = SaveRun(ctx, saved)) != XRT_OK) return rc;
  const bool wide = ContactsActive(ctx) || ParticlesOn(ctx) || BindingsOn(ctx);
  std::vector<xrt_tuning> cand;
  const xrt_tuning base{1024, 0, 1, wide ? 1 : 0, 0.0f, 0.0f, 0};
  if (wide) {
    for (int pt : {1, 2}) {
      xrt_tuning t = base;
      t.wide_per_thread = pt;
      cand.push_back(t);
    }
  } else {
    for (int block : {1024, 768, 512, 256})
      for (int unroll : {0, 1, 2}) {
        xrt_tuning t = base;
        t.persistent_block = block;
        t.persistent_unroll = unroll;
        cand.push_back(t);
      }
    for (int pt : {1, 2}) {  // the wide substep, which splits large instances over blocks
      xrt_tuning t = base;
      t.path = 1;
      t.wide_per_thread = pt;
      cand.push_back(t);
    }
  }
  const xrt_tuning before = ctx->tuning;
  xrt_tuning best = base;
  float best_ms = 0.0f, default_ms = 0.0f;
  for (size_t j = 0; j < cand.size(); ++j) {
    ctx->tuning = cand[j];
    float ms = 0.0f;
    rc = TimeCandidate(ctx, saved, substeps, &ms);
    if (rc != XRT_OK) break;
    if (j == 0) default_ms = ms;
    if (j == 0 || ms < best_ms) {
      best_ms = ms;
      best = cand[j];
    }
  }
  if (rc != XRT_OK) {
    ctx->tuning = before;
    (void)RestoreRun(ctx, saved);
    return rc;
  }
  best.best_ms = best_ms;
  best.default_ms = default_ms;
  best.candidates = static_cast<int32_t>(cand.size());
  ctx->tuning = best;
  DestroyGraph(ctx);
  if ((rc = CheckPairOverflow(ctx)) != XRT_OK) return rc;
  if (out) *out = best;
  return XRT_OK;
}

int xrt_time_batched_substep(xrt_ctx* ctx, int32_t n, float* avg_ms) {
  XRT_REQUIRE(ctx && ctx->started, XRT_ERR_STATE, "xrt_start has not been called");
  XRT_REQUIRE(n > 0 && avg_ms, XRT_ERR_INVALID_ARG, "n must be > 0");
  float total = 0.0f;
  for (int i = 0; i < n; ++i) {
    XRT_CHECK_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    int rc = SubstepBatched(ctx);
    if (rc != XRT_OK) return rc;
    XRT_CHECK_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    XRT_CHECK_HIP(hipEventSynchronize(ctx->ev1));
    float ms = 0.0f;
    XRT_CHECK_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    total += ms;
  }
  *avg_ms = total / static_cast<float>(n);
  return CheckPairOverflow(ctx);
}

int xrt_time_frames(xrt_ctx* ctx, int32_t batched, int32_t nframes, float* ms) {
  XRT_REQUIRE(ctx && ctx->started, XRT_ERR_STATE, "xrt_start has not been called");
  XRT_REQUIRE(nframes > 0 && ms, XRT_ERR_INVALID_ARG, "nframes must be > 0");
  if (batched && (!ctx->graph_exec || ScheduleKey(ctx) != ctx->graph_key)) {
    int rc = BuildGraph(ctx);
    if (rc != XRT_OK) return rc;
  }
  XRT_CHECK_HIP(hipEventRecord(ctx->ev0, ctx->stream));
  int rc = batched ? xrt_frames_batched(ctx, nframes) : xrt_frames_plain(ctx, nframes);
  if (rc != XRT_OK) return rc;
  XRT_CHECK_HIP(hipEventRecord(ctx->ev1, ctx->stream));
  XRT_CHECK_HIP(hipEventSynchronize(ctx->ev1));
  XRT_CHECK_HIP(hipEventElapsedTime(ms, ctx->ev0, ctx->ev1));
  return CheckPairOverflow(ctx);
}

}  // extern "C"
