// common.hpp -- Morton keys, status handling and small helpers shared by the host code.
//
// Key format (octree/src/morton.rs:1-140, 289-330): for each of 16 levels one bit triple
// (x, y, z), x highest, the triple of level 1 highest; the 48 position bits shifted left
// by 15 and the level in the low 15 bits.  Sorting keys as u64 gives Morton order with
// every ancestor before its descendants.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace fmmh {

constexpr int kDeepest = 16;
constexpr int kLevelBits = 15;
constexpr uint64_t kLevelMask = 0x7FFF;
constexpr uint32_t kNone = 0xFFFFFFFFu;

// Status codes of the C ABI; they mirror KernelError (fmm-kernels/src/error.rs:12-72).
enum Status : int {
  kOk = 0,
  kNotCompiled = 1,
  kNoDevice = 2,
  kUnsupportedPrecision = 3,
  kOutOfMemory = 4,
  kTooLarge = 5,
  kWrongDevice = 6,
  kUnsupportedLayout = 7,
  kDeviceError = 8,
  kInvalidArgument = 9,
};

struct Error : std::runtime_error {
  Status status;
  Error(Status s, const std::string &m) : std::runtime_error(m), status(s) {}
};

inline void require(bool ok, const std::string &what, Status s = kInvalidArgument) {
  if (!ok) throw Error(s, what);
}

// Spread the low 16 bits of v so bit b lands on bit 3b.
inline uint64_t spread3(uint64_t v) {
  v &= 0xFFFFULL;
  v = (v | (v << 16)) & 0x0000FF0000FFULL;
  v = (v | (v << 8)) & 0x00F00F00F00FULL;
  v = (v | (v << 4)) & 0x0C30C30C30C3ULL;
  v = (v | (v << 2)) & 0x249249249249ULL;
  return v;
}

inline uint64_t compact3(uint64_t v) {
  v &= 0x249249249249ULL;
  v = (v | (v >> 2)) & 0x0C30C30C30C3ULL;
  v = (v | (v >> 4)) & 0x00F00F00F00FULL;
  v = (v | (v >> 8)) & 0x0000FF0000FFULL;
  v = (v | (v >> 16)) & 0xFFFFULL;
  return v;
}

inline uint64_t make_key(int level, uint64_t ix, uint64_t iy, uint64_t iz) {
  const int s = kDeepest - level;
  const uint64_t pos = (spread3(ix << s) << 2) | (spread3(iy << s) << 1) | spread3(iz << s);
  return (pos << kLevelBits) | static_cast<uint64_t>(level);
}

inline int key_level(uint64_t k) { return static_cast<int>(k & kLevelMask); }

struct Cell {
  int level;
  int64_t i, j, k;
};

inline Cell decode(uint64_t key) {
  const int level = key_level(key);
  const uint64_t pos = key >> kLevelBits;
  const int s = kDeepest - level;
  return {level, static_cast<int64_t>(compact3(pos >> 2) >> s),
          static_cast<int64_t>(compact3(pos >> 1) >> s), static_cast<int64_t>(compact3(pos) >> s)};
}

inline uint64_t parent_key(uint64_t key) {
  const Cell c = decode(key);
  return make_key(c.level - 1, c.i >> 1, c.j >> 1, c.k >> 1);
}

// Child of octant o = 4x + 2y + z (CONVENTIONS §3.12 "Child octants").
inline uint64_t child_key(uint64_t key, int o) {
  const Cell c = decode(key);
  return make_key(c.level + 1, 2 * c.i + ((o >> 2) & 1), 2 * c.j + ((o >> 1) & 1),
                  2 * c.k + (o & 1));
}

inline int octant_of(uint64_t key) {
  const Cell c = decode(key);
  return static_cast<int>(((c.i & 1) << 2) | ((c.j & 1) << 1) | (c.k & 1));
}

inline uint64_t ancestor_at(uint64_t key, int level) {
  const Cell c = decode(key);
  const int s = c.level - level;
  return make_key(level, c.i >> s, c.j >> s, c.k >> s);
}

inline bool is_ancestor_or_self(uint64_t a, uint64_t b) {
  if (a == b) return true;
  const int la = key_level(a), lb = key_level(b);
  if (la > lb) return false;
  const int shift = kLevelBits + 3 * (kDeepest - la);
  return (a >> shift) == (b >> shift);
}

// Adjacency of fmm-plan/src/interaction_manager.rs:10-13: closed cubes intersect and
// neither box contains the other.
inline bool adjacent(uint64_t a, uint64_t b) {
  if (is_ancestor_or_self(a, b) || is_ancestor_or_self(b, a)) return false;
  const Cell ca = decode(a), cb = decode(b);
  const int L = ca.level > cb.level ? ca.level : cb.level;
  const int64_t sa = int64_t(1) << (L - ca.level), sb = int64_t(1) << (L - cb.level);
  const int64_t ia[3] = {ca.i, ca.j, ca.k}, ib[3] = {cb.i, cb.j, cb.k};
  for (int d = 0; d < 3; ++d) {
    const int64_t alo = ia[d] * sa, ahi = alo + sa, blo = ib[d] * sb, bhi = blo + sb;
    if (alo > bhi || blo > ahi) return false;
  }
  return true;
}

// Offset index of V_LIST_DIRECTIONS (fmm-plan/src/interaction_manager.rs:86), closed form
// of CONVENTIONS §3.12 "V-list offsets"; -1 when d is not one of the 316.
inline int offset_index(int dx, int dy, int dz) {
  auto inside = [](int t) { return t >= -3 && t <= 3; };
  auto small = [](int t) { return t >= -1 && t <= 1; };
  if (!inside(dx) || !inside(dy) || !inside(dz)) return -1;
  if (small(dx) && small(dy) && small(dz)) return -1;
  auto kappa = [](int t) { int v = t + 1; return v < 0 ? 0 : (v > 3 ? 3 : v); };
  int v = 49 * (dx + 3) + 7 * (dy + 3) + (dz + 3) - 9 * kappa(dx);
  if (small(dx)) v -= 3 * kappa(dy) + (small(dy) ? kappa(dz) : 0);
  return v;
}

inline void offset_of_index(int index, int d[3]) {
  int count = 0;
  for (int x = -3; x <= 3; ++x)
    for (int y = -3; y <= 3; ++y)
      for (int z = -3; z <= 3; ++z) {
        if (x >= -1 && x <= 1 && y >= -1 && y <= 1 && z >= -1 && z <= 1) continue;
        if (count++ == index) {
          d[0] = x;
          d[1] = y;
          d[2] = z;
          return;
        }
      }
}

inline int ncoef(int p) { return (p + 1) * (p + 1); }

}  // namespace fmmh
