// Swift-mode particle self-collision: set-up, scheduling and the C entry points.
//
// Schedule (ClothSolverBase::SimulateSwift, xrtailor/src/pipeline/base/ClothSolverBase.cpp:
// 190-203): after predict, SpatialHashGPU::Hash when the substep's index in its frame is a
// multiple of inter_leaved_hash; at the start of every iteration CollideParticles, then
// ApplyDeltas x relaxation_factor.  The initial positions (SpatialHashGPU::
// SetInitialPositions, SpatialHashGPU.cu:177-188) are the state at xrt_start.
#include <algorithm>
#include <cmath>
#include <vector>

#include "bvh_device.h"
#include "context.h"
#include "particle.h"

namespace xrt {

namespace {
constexpr int kVertBlock = 256;

ParticleBufs Bufs(xrt_ctx* c) {
  ParticleData& P = c->particles;
  ParticleBufs b;
  b.inst = P.dev.ptr;
  b.keys = P.keys.ptr;
  b.keys_tmp = P.keys_tmp.ptr;
  b.cell_start = P.cell_start.ptr;
  b.cell_end = P.cell_end.ptr;
  b.nb = P.nb.ptr;
  b.hist = P.hist.ptr;
  b.vals = P.vals.ptr;
  b.vals_tmp = P.vals_tmp.ptr;
  b.orig = P.orig.ptr;
  b.delta = P.delta.ptr;
  return b;
}

DevMap Map(const ParticleData& P, const MapRef& r) {
  DevMap m;
  m.off = P.maps.ptr + r.off;
  m.n = r.n;
  m.blocks = r.blocks;
  return m;
}

// particle diameter of an instance: its first triangle's first edge x factor
// (ClothObject.hpp:36-37), in FP32
float Diameter(const HostInstance& h, float factor) {
  const float* a = &h.rest_pos[3 * static_cast<size_t>(h.tris[0])];
  const float* b = &h.rest_pos[3 * static_cast<size_t>(h.tris[1])];
  const float dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
  return std::sqrt(dx * dx + dy * dy + dz * dz) * factor;
}

int CheckParticles(xrt_ctx* c, int32_t inst) {
  XRT_REQUIRE(c != nullptr, XRT_ERR_INVALID_ARG, "null context");
  XRT_REQUIRE(c->started, XRT_ERR_STATE, "xrt_start has not been called");
  XRT_REQUIRE(inst >= 0 && inst < static_cast<int32_t>(c->inst.size()), XRT_ERR_INVALID_ARG,
              "instance index out of range");
  XRT_REQUIRE(ParticlesOn(c) && c->particles.sized, XRT_ERR_STATE,
              "particle collision is off (enable_particle_collision)");
  return XRT_OK;
}

// the next free event pair of a timing run (nullptrs when none is running)
void NextPair(std::vector<hipEvent_t>& ev, size_t& next, hipEvent_t& a, hipEvent_t& b) {
  a = b = nullptr;
  if (next + 1 < ev.size()) {
    a = ev[next];
    b = ev[next + 1];
    next += 2;
  }
}

int Hash(xrt_ctx* c, int inst) {
  ParticleData& P = c->particles;
  const DevMap vm = Map(P, inst < 0 ? P.all_v : P.one_v[inst]);
  const DevMap tm = Map(P, inst < 0 ? P.all_t : P.one_t[inst]);
  hipEvent_t e0, e1;
  NextPair(P.hash_events, P.hash_next, e0, e1);
  if (e0) XRT_CHECK_HIP(hipEventRecord(e0, c->stream));
  XRT_CHECK_HIP(LaunchParticleHash(c->stream, Bufs(c), vm, tm, P.passes, c->x.ptr));
  if (e1) XRT_CHECK_HIP(hipEventRecord(e1, c->stream));
  return XRT_OK;
}

int Collide(xrt_ctx* c, int inst) {
  ParticleData& P = c->particles;
  const DevMap vm = Map(P, inst < 0 ? P.all_v : P.one_v[inst]);
  hipEvent_t e0, e1;
  NextPair(P.collide_events, P.collide_next, e0, e1);
  XRT_CHECK_HIP(LaunchCollideParticles(c->stream, Bufs(c), vm, c->x.ptr, c->x0.ptr,
                                       c->params.ground_friction, c->params.relaxation_factor,
                                       e0, e1));
  return XRT_OK;
}

// average elapsed ms over the recorded pairs [0, used)
int AverageMs(const std::vector<hipEvent_t>& ev, size_t used, float* avg) {
  float total = 0.0f;
  for (size_t q = 0; q + 1 < used; q += 2) {
    float ms = 0.0f;
    XRT_CHECK_HIP(hipEventElapsedTime(&ms, ev[q], ev[q + 1]));
    total += ms;
  }
  *avg = used >= 2 ? total / static_cast<float>(used / 2) : 0.0f;
  return XRT_OK;
}
}  // namespace

bool ParticlesOn(const xrt_ctx* ctx) { return ctx->params.enable_particle_collision != 0; }

int ParticleStart(xrt_ctx* ctx) {
  ParticleData& P = ctx->particles;
  const int n = static_cast<int>(ctx->inst.size());
  const int N = ctx->wide.n_total;
  P.substep.assign(n, 0);
  XRT_CHECK_HIP(P.orig.Alloc(std::max(N, 1)));
  XRT_CHECK_HIP(hipMemcpyAsync(P.orig.ptr, ctx->x0.ptr, sizeof(float4) * N, hipMemcpyDeviceToDevice,
                               ctx->stream));
  // block maps: every instance (per-vertex blocks, radix tiles), then each instance alone
  std::vector<int> vb(n), tb(n);
  for (int k = 0; k < n; ++k) {
    const int nv = ctx->inst[k].nv;
    vb[k] = std::max(1, (nv + kVertBlock - 1) / kVertBlock);
    tb[k] = std::max(1, (nv + kSortTile - 1) / kSortTile);
  }
  std::vector<int32_t> maps;
  P.all_v = AppendBlockMap(maps, 0, n, vb);
  P.all_t = AppendBlockMap(maps, 0, n, tb);
  P.one_v.resize(n);
  P.one_t.resize(n);
  for (int k = 0; k < n; ++k) {
    P.one_v[k] = AppendBlockMap(maps, k, k + 1, vb);
    P.one_t[k] = AppendBlockMap(maps, k, k + 1, tb);
  }
  XRT_CHECK_HIP(P.maps.Upload(maps, ctx->stream));
  XRT_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  return ParticleConfigure(ctx);
}

// descriptors from the current parameters; buffers (re)allocated when max_num_neighbors
// changes.  The neighbor lists start zeroed, like the reference's value-initialised vector.
int ParticleConfigure(xrt_ctx* ctx) {
  if (!ParticlesOn(ctx)) return XRT_OK;
  ParticleData& P = ctx->particles;
  const xrt_params& p = ctx->params;
  const int n = static_cast<int>(ctx->inst.size());
  const size_t N = static_cast<size_t>(ctx->wide.n_total);
  P.inst.assign(n, ParticleInstDev{});
  int64_t hsum = 0;
  int bits = 1;
  for (int k = 0; k < n; ++k) {
    const HostInstance& h = ctx->inst[k];
    ParticleInstDev& d = P.inst[k];
    d.vbeg = h.desc.vbeg;
    d.nv = h.nv;
    d.table = 2 * h.nv;
    d.max_nb = p.max_num_neighbors;
    d.nblocks = std::max(1, (h.nv + kSortTile - 1) / kSortTile);
    d.nbeg = static_cast<int64_t>(d.vbeg) * d.max_nb;
    d.hbeg = hsum;
    hsum += 256 * static_cast<int64_t>(d.nblocks);
    d.diameter = Diameter(h, p.particle_diameter);   // HashParams (SpatialHashGPU.cu:158-167)
    d.spacing = d.diameter * p.hash_cell_size;
    d.spacing2 = d.spacing * d.spacing;
    d.diameter2 = d.diameter * d.diameter;
    XRT_REQUIRE(std::isfinite(d.spacing2) && d.spacing > 0.0f, XRT_ERR_INVALID_ARG,
                "degenerate first edge: the particle diameter is zero");
    while ((int64_t{1} << bits) < d.table) ++bits;
  }
  // 8-bit digits, an even number of passes so that the sorted pairs end in (keys, vals)
  P.passes = (bits + 7) / 8;
  P.passes += P.passes & 1;
  XRT_CHECK_HIP(P.dev.Upload(P.inst, ctx->stream));
  if (!P.sized || P.max_nb != p.max_num_neighbors) {
    const size_t nb = N * static_cast<size_t>(p.max_num_neighbors);
    XRT_CHECK_HIP(P.keys.Alloc(std::max<size_t>(N, 1)));
    XRT_CHECK_HIP(P.keys_tmp.Alloc(std::max<size_t>(N, 1)));
    XRT_CHECK_HIP(P.vals.Alloc(std::max<size_t>(N, 1)));
    XRT_CHECK_HIP(P.vals_tmp.Alloc(std::max<size_t>(N, 1)));
    XRT_CHECK_HIP(P.cell_start.Alloc(std::max<size_t>(2 * N, 1)));
    XRT_CHECK_HIP(P.cell_end.Alloc(std::max<size_t>(2 * N, 1)));
    XRT_CHECK_HIP(P.nb.Alloc(std::max<size_t>(nb, 1)));
    XRT_CHECK_HIP(P.hist.Alloc(static_cast<size_t>(std::max<int64_t>(hsum, 1))));
    XRT_CHECK_HIP(P.delta.Alloc(std::max<size_t>(N, 1)));
    XRT_CHECK_HIP(hipMemsetAsync(P.nb.ptr, 0, sizeof(uint32_t) * std::max<size_t>(nb, 1), ctx->stream));
    XRT_CHECK_HIP(hipMemsetAsync(P.cell_start.ptr, 0xff, sizeof(uint32_t) * std::max<size_t>(2 * N, 1),
                                 ctx->stream));
    XRT_CHECK_HIP(hipMemsetAsync(P.cell_end.ptr, 0, sizeof(uint32_t) * std::max<size_t>(2 * N, 1),
                                 ctx->stream));
    P.max_nb = p.max_num_neighbors;
    P.sized = true;
  }
  XRT_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  return XRT_OK;
}

int ParticleSubstepBegin(xrt_ctx* ctx, int inst) {
  if (!ParticlesOn(ctx)) return XRT_OK;
  ParticleData& P = ctx->particles;
  XRT_REQUIRE(P.sized, XRT_ERR_STATE, "particle buffers missing");
  int sub = inst < 0 ? P.substep[0] : P.substep[inst];
  if (inst < 0)
    for (int s : P.substep)
      XRT_REQUIRE(s == sub, XRT_ERR_STATE,
                  "the batched substep needs every instance at the same substep of its frame");
  if (sub % ctx->params.inter_leaved_hash != 0) return XRT_OK;
  return Hash(ctx, inst);
}

int ParticleIter(xrt_ctx* ctx, int inst) {
  if (!ParticlesOn(ctx)) return XRT_OK;
  return Collide(ctx, inst);
}

int ParticleStateReset(xrt_ctx* ctx, int inst) {
  if (!ParticlesOn(ctx) || !ctx->particles.sized) return XRT_OK;
  XRT_REQUIRE(inst >= 0 && static_cast<size_t>(inst) < ctx->particles.one_v.size(), XRT_ERR_STATE,
              "particle maps missing for the instance");
  const int rc = Hash(ctx, inst);
  if (rc != XRT_OK) return rc;
  XRT_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  return XRT_OK;
}

int ParticleSubstepEnd(xrt_ctx* ctx, int inst) {
  ParticleData& P = ctx->particles;
  const int S = ctx->params.num_substeps;
  if (inst < 0) {
    for (int& s : P.substep) s = (s + 1) % S;
  } else if (inst < static_cast<int>(P.substep.size())) {
    P.substep[inst] = (P.substep[inst] + 1) % S;
  }
  return XRT_OK;
}

}  // namespace xrt

using namespace xrt;

extern "C" {

int xrt_particle_hash(xrt_ctx* ctx, int32_t inst) {
  int rc = CheckParticles(ctx, inst);
  if (rc != XRT_OK) return rc;
  return Hash(ctx, inst);
}

int xrt_collide_particles(xrt_ctx* ctx, int32_t inst) {
  int rc = CheckParticles(ctx, inst);
  if (rc != XRT_OK) return rc;
  return Collide(ctx, inst);
}

int xrt_time_particles(xrt_ctx* ctx, int32_t n, float ms[2], int64_t counts[2]) {
  XRT_REQUIRE(ctx && ctx->started, XRT_ERR_STATE, "xrt_start has not been called");
  XRT_REQUIRE(n > 0 && ms && counts, XRT_ERR_INVALID_ARG, "n must be > 0, outputs non-null");
  XRT_REQUIRE(ParticlesOn(ctx) && ctx->particles.sized, XRT_ERR_STATE,
              "particle collision is off (enable_particle_collision)");
  ParticleData& P = ctx->particles;
  XRT_REQUIRE(P.collide_events.empty() && P.hash_events.empty(), XRT_ERR_STATE,
              "particle timing already running");
  const size_t iters = static_cast<size_t>(std::max(1, ctx->params.num_iterations));
  P.collide_events.assign(2 * static_cast<size_t>(n) * iters, nullptr);
  P.hash_events.assign(2 * static_cast<size_t>(n), nullptr);
  P.collide_next = P.hash_next = 0;
  int rc = XRT_OK;
  for (auto* v : {&P.collide_events, &P.hash_events})
    for (auto& e : *v)
      if (hipEventCreate(&e) != hipSuccess) rc = XRT_ERR_HIP;
  for (int sub = 0; sub < n && rc == XRT_OK; ++sub) rc = xrt_substep_batched(ctx);
  if (rc == XRT_OK && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = XRT_ERR_HIP;
  if (rc == XRT_OK) rc = AverageMs(P.collide_events, P.collide_next, &ms[0]);
  if (rc == XRT_OK) rc = AverageMs(P.hash_events, P.hash_next, &ms[1]);
  const bool timed = P.collide_next >= 2;
  for (auto* v : {&P.collide_events, &P.hash_events}) {
    for (auto& e : *v)
      if (e) (void)hipEventDestroy(e);
    v->clear();
  }
  P.collide_next = P.hash_next = 0;
  if (rc != XRT_OK) return rc;
  XRT_REQUIRE(timed, XRT_ERR_STATE, "no collision launch was timed");
  // list entries of the last hash over every instance, and the particle count
  if (!P.count.ptr) XRT_CHECK_HIP(P.count.Alloc(1));
  XRT_CHECK_HIP(hipMemsetAsync(P.count.ptr, 0, sizeof(unsigned long long), ctx->stream));
  XRT_CHECK_HIP(LaunchCountNeighbors(ctx->stream, Bufs(ctx), Map(P, P.all_v), P.count.ptr));
  unsigned long long entries = 0;
  XRT_CHECK_HIP(hipMemcpyAsync(&entries, P.count.ptr, sizeof(entries), hipMemcpyDeviceToHost,
                               ctx->stream));
  XRT_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  counts[0] = static_cast<int64_t>(entries);
  counts[1] = ctx->wide.n_total;
  return XRT_OK;
}

int xrt_get_particle_hash(xrt_ctx* ctx, int32_t inst, uint32_t* neighbors, uint32_t* sorted_hash,
                          int32_t* sorted_index, uint32_t* cell_start, uint32_t* cell_end,
                          float info[2]) {
  int rc = CheckParticles(ctx, inst);
  if (rc != XRT_OK) return rc;
  const ParticleData& P = ctx->particles;
  const ParticleInstDev& d = P.inst[inst];
  hipStream_t s = ctx->stream;
  const size_t nv = static_cast<size_t>(d.nv);
  if (neighbors)
    XRT_CHECK_HIP(hipMemcpyAsync(neighbors, P.nb.ptr + d.nbeg, sizeof(uint32_t) * nv * d.max_nb,
                                 hipMemcpyDeviceToHost, s));
  if (sorted_hash)
    XRT_CHECK_HIP(hipMemcpyAsync(sorted_hash, P.keys.ptr + d.vbeg, sizeof(uint32_t) * nv,
                                 hipMemcpyDeviceToHost, s));
  if (sorted_index)
    XRT_CHECK_HIP(hipMemcpyAsync(sorted_index, P.vals.ptr + d.vbeg, sizeof(int32_t) * nv,
                                 hipMemcpyDeviceToHost, s));
  if (cell_start)
    XRT_CHECK_HIP(hipMemcpyAsync(cell_start, P.cell_start.ptr + 2 * static_cast<size_t>(d.vbeg),
                                 sizeof(uint32_t) * 2 * nv, hipMemcpyDeviceToHost, s));
  if (cell_end)
    XRT_CHECK_HIP(hipMemcpyAsync(cell_end, P.cell_end.ptr + 2 * static_cast<size_t>(d.vbeg),
                                 sizeof(uint32_t) * 2 * nv, hipMemcpyDeviceToHost, s));
  if (info) {
    info[0] = d.diameter;
    info[1] = d.spacing;
  }
  XRT_CHECK_HIP(hipStreamSynchronize(s));
  return XRT_OK;
}

}  // extern "C"
