// Host side of the multi-instance block maps (device side: XcdBlock, InstOfBlock and
// InstOfEntry in device.h).  A map over n entries holds the n + 1 prefix sums of their
// block counts, the entry (0..n-1) of every block, then the instance of every entry, so
// one launch covers every instance (the all-instance map) or a single one.
#pragma once
#include <cstddef>
#include <cstdint>
#include <vector>

namespace xrt {

struct MapRef {
  size_t off = 0;  // offset of the map in its device array
  int n = 0, blocks = 0;
};

// a map on the device (off: n + 1 prefix sums over its entries, the entry of every block,
// the instance of every entry) and its block count
struct DevMap {
  const int32_t* off = nullptr;
  int n = 0, blocks = 0;
};

// appends the map of instances [k0, k1) with per-instance block counts `blocks`
inline MapRef AppendBlockMap(std::vector<int32_t>& maps, int k0, int k1,
                             const std::vector<int>& blocks) {
  MapRef r;
  r.off = maps.size();
  r.n = k1 - k0;
  std::vector<int32_t> o(r.n + 1, 0);
  for (int j = 0; j < r.n; ++j) o[j + 1] = o[j] + blocks[k0 + j];
  r.blocks = o[r.n];
  maps.insert(maps.end(), o.begin(), o.end());
  for (int j = 0; j < r.n; ++j) maps.insert(maps.end(), static_cast<size_t>(blocks[k0 + j]), j);
  for (int j = 0; j < r.n; ++j) maps.push_back(k0 + j);
  return r;
}

}  // namespace xrt
