/*
 * particle_oracle.c -- serial restatement of XRTailor's swift-mode particle self-collision:
 * the spatial hash (SpatialHashGPU::Hash, xrtailor/src/physics/broad_phase/spatial_hash/
 * SpatialHashGPU.cu:9-176) and CollideParticles (xrtailor/src/physics/ClothSolverHelper.cu:
 * 156-219).
 *
 * TEST INFRASTRUCTURE ONLY (see xrt_oracle.h).
 *
 * The hash and the neighbor search produce integer outputs from FP32 positions and are
 * restated in FP32 (this file is built with -ffp-contract=off, so every float operation
 * rounds once, as in the strict-FP device code); the collision response is FP64.
 *
 * Pinned by (tests/test_oracle_particle.py): HashCoords against hand-computed values
 * (wrapping 32-bit products, C remainder sign), the neighbor lists against a brute-force
 * restatement that scans every particle per visited cell (no sort, no cell table), per-cell
 * truncation and the full-list early return at small max_num_neighbors, and the response
 * against two-particle closed forms (separation to exactly the diameter at relaxation 1,
 * Coulomb friction of a tangential relative velocity, pinned partners, the count).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "xrt_oracle.h"

#define EPS 1e-6

/* HashCoords (SpatialHashGPU.cu:13-16): 32-bit wrapping products, C remainder, abs */
int32_t orc_hash_coords(int32_t x, int32_t y, int32_t z, int32_t table) {
  const uint32_t h = ((uint32_t)x * 92837111u) ^ ((uint32_t)y * 689287499u) ^
                     ((uint32_t)z * 283923481u);
  const int32_t r = (int32_t)h % table;
  return r < 0 ? -r : r;
}

/* ComputeIntCoord (SpatialHashGPU.cu:9-11): floor of the FP32 quotient */
static int32_t int_coord(float v, float spacing) { return (int32_t)floorf(v / spacing); }

static float len2f(float dx, float dy, float dz) { return dx * dx + dy * dy + dz * dz; }

typedef struct { uint32_t h; uint32_t id; } hkey;

static int cmp_hkey(const void* a, const void* b) {
  const hkey* p = (const hkey*)a;
  const hkey* q = (const hkey*)b;
  if (p->h != q->h) return p->h < q->h ? -1 : 1;
  return p->id < q->id ? -1 : (p->id > q->id);
}

/* SpatialHashGPU::Hash over n particles at FP32 positions x (3n) with FP32 initial
 * positions orig (3n): table 2n, cell spacing d * cell_size, neighbor lists column-major
 * (entry k of particle i at nb[k * n + i]) ended by 0xffffffff unless full.  The sort is
 * the reference's stable radix sort of (hash, index): ties keep ascending index.  Optional
 * outputs (null to skip): the sorted hashes and indices (n) and the cell table (2n; cell_end
 * is meaningful only where cell_start is set). */
void orc_particle_neighbors(const float* x, const float* orig, int n, float diameter,
                            float cell_size, int max_nb, uint32_t* nb, uint32_t* s_hash,
                            uint32_t* s_index, uint32_t* cell_start, uint32_t* cell_end) {
  const float spacing = diameter * cell_size;
  const float spacing2 = spacing * spacing;
  const float d2 = diameter * diameter;
  const int32_t table = 2 * n;
  hkey* k = (hkey*)malloc(sizeof(hkey) * (size_t)(n > 0 ? n : 1));
  uint32_t* cs = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(table > 0 ? table : 1));
  uint32_t* ce = (uint32_t*)calloc((size_t)(table > 0 ? table : 1), sizeof(uint32_t));
  for (int i = 0; i < n; i++) {  /* ComputeParticleHash_Kernel */
    k[i].h = (uint32_t)orc_hash_coords(int_coord(x[3 * i], spacing), int_coord(x[3 * i + 1], spacing),
                                       int_coord(x[3 * i + 2], spacing), table);
    k[i].id = (uint32_t)i;
  }
  qsort(k, (size_t)n, sizeof(hkey), cmp_hkey);
  for (int c = 0; c < table; c++) cs[c] = 0xffffffffu;
  for (int t = 0; t < n; t++) {  /* FindCellStart_Kernel */
    if (t == 0 || k[t].h != k[t - 1].h) {
      cs[k[t].h] = (uint32_t)t;
      if (t > 0) ce[k[t - 1].h] = (uint32_t)t;
    }
    if (t == n - 1) ce[k[t].h] = (uint32_t)n;
  }
  const long cap = (long)n * max_nb;
  for (int t = 0; t < n; t++) {  /* CacheNeighbors_Kernel, in sorted order */
    const uint32_t id = k[t].id;
    const float px = x[3 * id], py = x[3 * id + 1], pz = x[3 * id + 2];
    const float ox = orig[3 * id], oy = orig[3 * id + 1], oz = orig[3 * id + 2];
    const int32_t ix = int_coord(px, spacing), iy = int_coord(py, spacing), iz = int_coord(pz, spacing);
    long at = id;
    int full = 0;
    for (int32_t cx = ix - 1; cx <= ix + 1 && !full; cx++)
      for (int32_t cy = iy - 1; cy <= iy + 1 && !full; cy++)
        for (int32_t cz = iz - 1; cz <= iz + 1 && !full; cz++) {
          const int32_t h = orc_hash_coords(cx, cy, cz, table);
          const uint32_t start = cs[h];
          if (start == 0xffffffffu) continue;
          const uint32_t end = ce[h] < start + (uint32_t)max_nb ? ce[h] : start + (uint32_t)max_nb;
          for (uint32_t s = start; s < end && !full; s++) {
            const uint32_t j = k[s].id;
            if (j == id) continue;
            if (!(len2f(px - x[3 * j], py - x[3 * j + 1], pz - x[3 * j + 2]) < spacing2)) continue;
            if (!(len2f(ox - orig[3 * j], oy - orig[3 * j + 1], oz - orig[3 * j + 2]) > d2)) continue;
            nb[at] = j;
            at += n;
            if (at >= cap) full = 1;
          }
        }
    if (!full && at < cap) nb[at] = 0xffffffffu;
  }
  for (int t = 0; t < n; t++) {
    if (s_hash) s_hash[t] = k[t].h;
    if (s_index) s_index[t] = k[t].id;
  }
  if (cell_start) memcpy(cell_start, cs, sizeof(uint32_t) * (size_t)table);
  if (cell_end) memcpy(cell_end, ce, sizeof(uint32_t) * (size_t)table);
  free(k);
  free(cs);
  free(ce);
}

/* CollideParticles_Kernel (ClothSolverHelper.cu:171-211) with ComputeFriction (:156-169):
 * delta (3n) and count (n) are assigned, not accumulated; x, x0 FP64 (3n), w inverse
 * masses, friction mu (SimParams::sdf_friction). */
void orc_collide_particles(const double* x, const double* x0, const double* w, int n,
                           const uint32_t* nb, int max_nb, double diameter, double mu,
                           double* delta, int32_t* count) {
  for (int i = 0; i < n; i++) {
    double d[3] = {0, 0, 0};
    int cnt = 0;
    const double vi[3] = {x[3 * i] - x0[3 * i], x[3 * i + 1] - x0[3 * i + 1], x[3 * i + 2] - x0[3 * i + 2]};
    for (long s = i; s < (long)n * max_nb; s += n) {
      const uint32_t j = nb[s];
      if (j > (uint32_t)n) break;
      const double denom = w[i] + w[j];
      if (denom <= 0) continue;
      double diff[3];
      for (int a = 0; a < 3; a++) diff[a] = x[3 * i + a] - x[3 * j + a];
      const double dist = sqrt(diff[0] * diff[0] + diff[1] * diff[1] + diff[2] * diff[2]);
      if (dist >= diameter) continue;
      const double lambda = (dist - diameter) / denom;
      double common[3], rel[3];
      for (int a = 0; a < 3; a++) {
        common[a] = lambda * (diff[a] / (dist + EPS));
        rel[a] = vi[a] - (x[3 * j + a] - x0[3 * j + a]);
      }
      cnt++;
      for (int a = 0; a < 3; a++) d[a] -= w[i] * common[a];
      const double cl = sqrt(common[0] * common[0] + common[1] * common[1] + common[2] * common[2]);
      if (mu > 0 && cl > 0) {
        double nrm[3], tan[3];
        for (int a = 0; a < 3; a++) nrm[a] = common[a] / cl;
        const double vn = rel[0] * nrm[0] + rel[1] * nrm[1] + rel[2] * nrm[2];
        for (int a = 0; a < 3; a++) tan[a] = rel[a] - nrm[a] * vn;
        const double tl = sqrt(tan[0] * tan[0] + tan[1] * tan[1] + tan[2] * tan[2]);
        const double s = tl > 0 ? fmin(cl * mu / tl, 1.0) : 1.0;
        for (int a = 0; a < 3; a++) d[a] += w[i] * (-tan[a] * s);
      }
    }
    for (int a = 0; a < 3; a++) delta[3 * i + a] = d[a];
    count[i] = cnt;
  }
}
