// C ABI of the standalone LBVH (the reference's BVH class surface and Traverse).
#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

#include "bvh.h"
#include "context.h"

struct xrt_bvh {
  int device = 0;
  hipStream_t stream = nullptr;
  xrt::Lbvh tree;
  xrt::DevBuf<float4> a, b, c;  // staged per-vertex inputs
  xrt::DevBuf<int32_t> vptr, corner;
};

// Device staging of xrt_bvh_traverse_noalloc's pairs: a fixed array in place of the
// per-call allocation of xrt_bvh_traverse.  Each device holds its own copy until the
// process ends; every call on one device shares it (calls that overlap on a device race).
// Nothing in it needs to start at zero: a call writes every slot it reads back.
constexpr int64_t kNoAllocMaxPairs = int64_t{1} << 21;  // 2,097,152 pairs, 16 MiB
static __device__ __align__(256) int2 scratch[kNoAllocMaxPairs];

namespace {
using xrt::DevBuf;

// device address of `scratch` per device, resolved once per device by xrt_bvh_create with
// that device current (the array's host-side name is not a device address)
std::mutex g_scratch_mu;
std::vector<int2*> g_scratch_ptr;

int ResolveScratch(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (static_cast<int>(g_scratch_ptr.size()) <= device) g_scratch_ptr.resize(device + 1, nullptr);
  if (g_scratch_ptr[device]) return XRT_OK;
  void* p = nullptr;
  XRT_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  XRT_REQUIRE(p != nullptr, XRT_ERR_HIP, "no device address for the pair staging array");
  g_scratch_ptr[device] = static_cast<int2*>(p);
  return XRT_OK;
}

int2* ScratchOf(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return device < static_cast<int>(g_scratch_ptr.size()) ? g_scratch_ptr[device] : nullptr;
}

int Stage(xrt_bvh* h, DevBuf<float4>& dst, const float* src, int n) {
  std::vector<float4> tmp(std::max(n, 1));
  for (int i = 0; i < n; ++i) tmp[i] = make_float4(src[3 * i], src[3 * i + 1], src[3 * i + 2], 0.0f);
  if (dst.n < tmp.size()) XRT_CHECK_HIP(dst.Alloc(tmp.size()));
  XRT_CHECK_HIP(dst.Upload(tmp.data(), tmp.size(), h->stream));
  return XRT_OK;
}

int CheckMesh(const float* x0, int nv, const int32_t* tris, int nt) {
  XRT_REQUIRE(x0 && tris && nv > 0 && nt >= 0, XRT_ERR_INVALID_ARG, "null or empty mesh");
  for (int i = 0; i < 3 * nt; ++i)
    XRT_REQUIRE(tris[i] >= 0 && tris[i] < nv, XRT_ERR_INVALID_ARG, "face index out of range");
  for (int i = 0; i < 3 * nv; ++i)
    XRT_REQUIRE(std::isfinite(x0[i]), XRT_ERR_NOT_FINITE, "non-finite vertex position");
  return XRT_OK;
}
}  // namespace

extern "C" {

int xrt_bvh_create(xrt_bvh** out, int device) {
  XRT_REQUIRE(out != nullptr, XRT_ERR_INVALID_ARG, "null output pointer");
  *out = nullptr;
  int ndev = 0;
  XRT_CHECK_HIP(hipGetDeviceCount(&ndev));
  XRT_REQUIRE(device >= 0 && device < ndev, XRT_ERR_INVALID_ARG, "device index out of range");
  XRT_CHECK_HIP(hipSetDevice(device));
  const int rc = ResolveScratch(device);  // device is current
  if (rc != XRT_OK) return rc;
  auto* h = new xrt_bvh();
  h->device = device;
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
    delete h;
    xrt::SetError("stream creation failed");
    return XRT_ERR_HIP;
  }
  *out = h;
  return XRT_OK;
}

int xrt_bvh_destroy(xrt_bvh* h) {
  if (!h) return XRT_OK;
  (void)hipSetDevice(h->device);
  (void)hipStreamSynchronize(h->stream);
  (void)hipStreamDestroy(h->stream);
  delete h;
  return XRT_OK;
}

int xrt_bvh_build(xrt_bvh* h, const float* x0, const float* x, const float* n, int32_t nv,
                  const int32_t* tris, int32_t nt, float tol, int32_t ccd, float scr) {
  XRT_REQUIRE(h, XRT_ERR_INVALID_ARG, "null bvh");
  int rc = CheckMesh(x0, nv, tris, nt);
  if (rc != XRT_OK) return rc;
  XRT_REQUIRE(!ccd || x, XRT_ERR_INVALID_ARG, "ccd needs predicted positions");
  XRT_CHECK_HIP(hipSetDevice(h->device));
  XRT_CHECK_HIP(h->tree.Setup(tris, nt, nv, h->stream));
  if ((rc = Stage(h, h->a, x0, nv)) != XRT_OK) return rc;
  if (x && (rc = Stage(h, h->b, x, nv)) != XRT_OK) return rc;
  if (n && (rc = Stage(h, h->c, n, nv)) != XRT_OK) return rc;
  XRT_CHECK_HIP(xrt::LbvhBuild(h->tree, h->a.ptr, x ? h->b.ptr : nullptr, n ? h->c.ptr : nullptr,
                               tol, ccd, scr, h->stream));
  XRT_CHECK_HIP(hipStreamSynchronize(h->stream));
  return XRT_OK;
}

int xrt_bvh_refit_midstep(xrt_bvh* h, const float* x1, const float* n1, float tol) {
  XRT_REQUIRE(h && x1, XRT_ERR_INVALID_ARG, "null argument");
  XRT_REQUIRE(h->tree.active, XRT_ERR_STATE, "bvh not built");
  int rc;
  if ((rc = Stage(h, h->a, x1, h->tree.nv)) != XRT_OK) return rc;
  if (n1 && (rc = Stage(h, h->c, n1, h->tree.nv)) != XRT_OK) return rc;
  XRT_CHECK_HIP(xrt::LbvhRefitMidstep(h->tree, h->a.ptr, n1 ? h->c.ptr : nullptr, tol, h->stream));
  XRT_CHECK_HIP(hipStreamSynchronize(h->stream));
  return XRT_OK;
}

int xrt_bvh_update(xrt_bvh* h, const float* x, const float* x0, int32_t ccd) {
  XRT_REQUIRE(h && x && (!ccd || x0), XRT_ERR_INVALID_ARG, "null argument");
  XRT_REQUIRE(h->tree.active, XRT_ERR_STATE, "bvh not built");
  int rc;
  if ((rc = Stage(h, h->a, x, h->tree.nv)) != XRT_OK) return rc;
  if (ccd && (rc = Stage(h, h->b, x0, h->tree.nv)) != XRT_OK) return rc;
  XRT_CHECK_HIP(xrt::LbvhUpdate(h->tree, h->a.ptr, ccd ? h->b.ptr : nullptr, ccd, h->stream));
  XRT_CHECK_HIP(hipStreamSynchronize(h->stream));
  return XRT_OK;
}

int xrt_bvh_num_objects(xrt_bvh* h, int32_t* n) {
  XRT_REQUIRE(h && n, XRT_ERR_INVALID_ARG, "null argument");
  *n = h->tree.active ? h->tree.nf : 0;
  return XRT_OK;
}

int xrt_bvh_get(xrt_bvh* h, uint32_t* morton, int32_t* order, int32_t* nodes, float* aabbs,
                int32_t* i_escape, int32_t* e_escape, float* root) {
  XRT_REQUIRE(h, XRT_ERR_INVALID_ARG, "null bvh");
  XRT_REQUIRE(h->tree.active, XRT_ERR_STATE, "bvh not built");
  const int nf = h->tree.nf, nn = 2 * nf - 1, ni = nf - 1;
  hipStream_t s = h->stream;
  if (morton) XRT_CHECK_HIP(hipMemcpyAsync(morton, h->tree.morton.ptr, 4 * nf, hipMemcpyDeviceToHost, s));
  if (order) XRT_CHECK_HIP(hipMemcpyAsync(order, h->tree.order.ptr, 4 * nf, hipMemcpyDeviceToHost, s));
  std::vector<int4> nd(nn);
  std::vector<xrt::Box> bx(nn);
  std::vector<int32_t> esc(nn);
  std::vector<uint32_t> rt(6);
  XRT_CHECK_HIP(hipMemcpyAsync(nd.data(), h->tree.nodes.ptr, sizeof(int4) * nn, hipMemcpyDeviceToHost, s));
  XRT_CHECK_HIP(hipMemcpyAsync(bx.data(), h->tree.box.ptr, sizeof(xrt::Box) * nn, hipMemcpyDeviceToHost, s));
  XRT_CHECK_HIP(hipMemcpyAsync(esc.data(), h->tree.escape.ptr, 4 * nn, hipMemcpyDeviceToHost, s));
  XRT_CHECK_HIP(hipMemcpyAsync(rt.data(), h->tree.root.ptr, 24, hipMemcpyDeviceToHost, s));
  XRT_CHECK_HIP(hipStreamSynchronize(s));
  for (int k = 0; k < nn; ++k) {
    if (nodes) {
      nodes[4 * k] = nd[k].x; nodes[4 * k + 1] = nd[k].y;
      nodes[4 * k + 2] = nd[k].z; nodes[4 * k + 3] = nd[k].w;
    }
    if (aabbs) {
      const float v[6] = {bx[k].lo.x, bx[k].lo.y, bx[k].lo.z, bx[k].hi.x, bx[k].hi.y, bx[k].hi.z};
      std::copy(v, v + 6, aabbs + 6 * k);
    }
  }
  // the reference's two tables: internal (size F, entry F-1 unused = -1) and external
  if (i_escape) {
    for (int k = 0; k < ni; ++k) i_escape[k] = esc[k];
    i_escape[nf - 1] = -1;
  }
  if (e_escape)
    for (int k = 0; k < nf; ++k) e_escape[k] = esc[ni + k];
  if (root)
    for (int k = 0; k < 6; ++k) root[k] = xrt::DecodeOrdered(rt[k]);
  return XRT_OK;
}

int xrt_bvh_query(xrt_bvh* h, const float* p0, const float* p1, int32_t n, int32_t* hits,
                  int32_t* counts) {
  XRT_REQUIRE(h && p0 && hits && counts && n >= 0, XRT_ERR_INVALID_ARG, "null argument");
  XRT_REQUIRE(h->tree.active, XRT_ERR_STATE, "bvh not built");
  if (n == 0) return XRT_OK;
  DevBuf<float4> q0, q1;
  std::vector<float4> tmp(n);
  for (int i = 0; i < n; ++i) tmp[i] = make_float4(p0[3 * i], p0[3 * i + 1], p0[3 * i + 2], 0.0f);
  XRT_CHECK_HIP(q0.Upload(tmp, h->stream));
  if (p1) {
    for (int i = 0; i < n; ++i) tmp[i] = make_float4(p1[3 * i], p1[3 * i + 1], p1[3 * i + 2], 0.0f);
    XRT_CHECK_HIP(q1.Upload(tmp, h->stream));
  }
  DevBuf<int32_t> dh, dc;
  XRT_CHECK_HIP(dh.Alloc(static_cast<size_t>(n) * 256));
  XRT_CHECK_HIP(dc.Alloc(n));
  XRT_CHECK_HIP(xrt::LbvhQuery(h->tree, q0.ptr, p1 ? q1.ptr : nullptr, n, dh.ptr, dc.ptr, h->stream));
  XRT_CHECK_HIP(hipMemcpyAsync(hits, dh.ptr, sizeof(int32_t) * 256 * n, hipMemcpyDeviceToHost, h->stream));
  XRT_CHECK_HIP(hipMemcpyAsync(counts, dc.ptr, sizeof(int32_t) * n, hipMemcpyDeviceToHost, h->stream));
  XRT_CHECK_HIP(hipStreamSynchronize(h->stream));
  return XRT_OK;
}

int xrt_bvh_traverse(xrt_bvh* lhs, xrt_bvh* rhs, float thickness, int32_t* pairs, int64_t cap,
                     int64_t* n_out) {
  XRT_REQUIRE(lhs && rhs && n_out && cap >= 0 && (cap == 0 || pairs), XRT_ERR_INVALID_ARG,
              "null argument");
  XRT_REQUIRE(lhs->device == rhs->device, XRT_ERR_INVALID_ARG, "trees on different devices");
  *n_out = 0;
  if (!lhs->tree.active || !rhs->tree.active) return XRT_OK;
  hipStream_t s = lhs->stream;
  XRT_CHECK_HIP(hipStreamSynchronize(rhs->stream));
  XRT_CHECK_HIP(xrt::LbvhCountPairs(lhs->tree, rhs->tree, thickness, s));
  int32_t total = 0;
  XRT_CHECK_HIP(hipMemcpyAsync(&total, lhs->tree.total.ptr, 4, hipMemcpyDeviceToHost, s));
  XRT_CHECK_HIP(hipStreamSynchronize(s));
  *n_out = total;
  const int64_t take = std::min<int64_t>(total, cap);
  if (take > 0) {
    DevBuf<int2> dp;
    XRT_CHECK_HIP(dp.Alloc(take));
    XRT_CHECK_HIP(xrt::LbvhFindPairs(lhs->tree, rhs->tree, thickness, dp.ptr, static_cast<int>(take), s));
    XRT_CHECK_HIP(hipMemcpyAsync(pairs, dp.ptr, sizeof(int2) * take, hipMemcpyDeviceToHost, s));
    XRT_CHECK_HIP(hipStreamSynchronize(s));
  }
  XRT_REQUIRE(total <= cap, XRT_ERR_CAPACITY, "pair buffer smaller than the pair count");
  return XRT_OK;
}

int xrt_bvh_traverse_noalloc(xrt_bvh* lhs, xrt_bvh* rhs, float thickness, int32_t* pairs,
                             int64_t cap, int64_t* n_out) {
  XRT_REQUIRE(lhs && rhs && n_out && cap >= 0 && (cap == 0 || pairs), XRT_ERR_INVALID_ARG,
              "null argument");
  XRT_REQUIRE(lhs->device == rhs->device, XRT_ERR_INVALID_ARG, "trees on different devices");
  *n_out = 0;
  if (!lhs->tree.active || !rhs->tree.active) return XRT_OK;
  int2* stage = ScratchOf(lhs->device);
  XRT_REQUIRE(stage != nullptr, XRT_ERR_STATE, "pair staging array not resolved for this device");
  hipStream_t s = lhs->stream;
  XRT_CHECK_HIP(hipStreamSynchronize(rhs->stream));
  // count pass into the trees' own per-leaf counts / offsets (allocated by xrt_bvh_build)
  XRT_CHECK_HIP(xrt::LbvhCountPairs(lhs->tree, rhs->tree, thickness, s));
  int32_t total = 0;
  XRT_CHECK_HIP(hipMemcpyAsync(&total, lhs->tree.total.ptr, 4, hipMemcpyDeviceToHost, s));
  XRT_CHECK_HIP(hipStreamSynchronize(s));
  *n_out = total;
  const int64_t take = std::min<int64_t>(total, cap);
  XRT_REQUIRE(take <= kNoAllocMaxPairs, XRT_ERR_CAPACITY,
              "pairs to return (" + std::to_string(take) + ") exceed the no-allocation staging "
              "maximum of 2,097,152 pairs; use xrt_bvh_traverse");
  if (take > 0) {
    XRT_CHECK_HIP(xrt::LbvhFindPairs(lhs->tree, rhs->tree, thickness, stage, static_cast<int>(take), s));
    XRT_CHECK_HIP(hipMemcpyAsync(pairs, stage, sizeof(int2) * take, hipMemcpyDeviceToHost, s));
    XRT_CHECK_HIP(hipStreamSynchronize(s));
  }
  XRT_REQUIRE(total <= cap, XRT_ERR_CAPACITY, "pair buffer smaller than the pair count");
  return XRT_OK;
}

int xrt_vertex_normals(xrt_bvh* h, const float* x, int32_t nv, const int32_t* tris, int32_t nt,
                       float* n_out) {
  XRT_REQUIRE(h && n_out, XRT_ERR_INVALID_ARG, "null argument");
  int rc = CheckMesh(x, nv, tris, nt);
  if (rc != XRT_OK) return rc;
  std::vector<int32_t> vptr(nv + 1, 0), corner(3 * nt);
  for (int i = 0; i < 3 * nt; ++i) vptr[tris[i] + 1]++;
  for (int v = 0; v < nv; ++v) vptr[v + 1] += vptr[v];
  std::vector<int32_t> fill(vptr.begin(), vptr.end() - 1);
  for (int i = 0; i < 3 * nt; ++i) corner[fill[tris[i]]++] = i;  // ascending face order
  std::vector<int4> t(std::max(nt, 1));
  for (int f = 0; f < nt; ++f) t[f] = make_int4(tris[3 * f], tris[3 * f + 1], tris[3 * f + 2], 0);
  DevBuf<int4> dt;
  DevBuf<float4> dn;
  XRT_CHECK_HIP(dt.Upload(t, h->stream));
  XRT_CHECK_HIP(h->vptr.Upload(vptr, h->stream));
  XRT_CHECK_HIP(h->corner.Upload(corner, h->stream));
  if ((rc = Stage(h, h->a, x, nv)) != XRT_OK) return rc;
  XRT_CHECK_HIP(dn.Alloc(nv));
  XRT_CHECK_HIP(xrt::VertexNormals(h->a.ptr, dt.ptr, h->vptr.ptr, h->corner.ptr, nv, dn.ptr, h->stream));
  std::vector<float4> out(nv);
  XRT_CHECK_HIP(hipMemcpyAsync(out.data(), dn.ptr, sizeof(float4) * nv, hipMemcpyDeviceToHost, h->stream));
  XRT_CHECK_HIP(hipStreamSynchronize(h->stream));
  for (int v = 0; v < nv; ++v) {
    n_out[3 * v] = out[v].x; n_out[3 * v + 1] = out[v].y; n_out[3 * v + 2] = out[v].z;
  }
  return XRT_OK;
}

}  // extern "C"
