// table_cache.cpp -- versioned on-disk cache of the Laplace operator tables (the
// reference's fmm-tables/src/cache.rs, its file format kept so that either side can read
// the other's files).
//
// One file per table family and degree in a directory the caller names
// (fmmh_set_table_cache): `<kind>-p<pp>-f64-c<convention>-v<format>.bin`, a 44-byte
// little-endian header -- magic "NDFMMTAB", format version, convention version, kind
// (1 M2M, 2 L2L, 3 M2L), degree, bits per value (64), payload length, FNV-1a 64 checksum
// of the payload -- then the matrices in their in-memory order (column-major, octant or
// offset order, CONVENTIONS section 3.12).  A store writes a uniquely named temporary
// file in the directory, syncs it and renames it over the final name, so no partial file
// appears under a final name and concurrent writers of one key both succeed.  A load
// checks the header length, magic, versions, key, payload length and checksum in that
// order; any failure, or a missing file, builds the tables and stores them (a load never
// fails because of the cache).
#include <fcntl.h>
#include <unistd.h>

#include <atomic>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "plan.hpp"
#include "table_cache.hpp"

namespace fmmh {

namespace {

constexpr char kMagic[8] = {'N', 'D', 'F', 'M', 'M', 'T', 'A', 'B'};
constexpr uint32_t kFormatVersion = 1;
// the reference's convention version of the tables' conventions (fmm-math/src/lib.rs:73)
constexpr uint32_t kConventionVersion = 4;
constexpr size_t kHeaderLen = 44;

std::mutex g_mu;
std::string g_dir;
int64_t g_hits = 0, g_misses = 0;
std::atomic<uint64_t> g_counter{0};

uint64_t fnv1a(const unsigned char *p, size_t n) {
  uint64_t h = 0xcbf29ce484222325ull;  // the FNV-1a 64-bit offset basis
  for (size_t i = 0; i < n; ++i) {
    h ^= p[i];
    h *= 0x100000001b3ull;  // the FNV 64-bit prime
  }
  return h;
}

void put32(unsigned char *b, uint32_t v) {
  for (int i = 0; i < 4; ++i) b[i] = static_cast<unsigned char>(v >> (8 * i));
}
void put64(unsigned char *b, uint64_t v) {
  for (int i = 0; i < 8; ++i) b[i] = static_cast<unsigned char>(v >> (8 * i));
}
uint32_t get32(const unsigned char *b) {
  uint32_t v = 0;
  for (int i = 3; i >= 0; --i) v = (v << 8) | b[i];
  return v;
}
uint64_t get64(const unsigned char *b) {
  uint64_t v = 0;
  for (int i = 7; i >= 0; --i) v = (v << 8) | b[i];
  return v;
}

const char *kind_name(int kind) { return kind == 1 ? "m2m" : kind == 2 ? "l2l" : "m2l"; }

std::string file_name(int kind, int p) {
  char buf[64];
  std::snprintf(buf, sizeof buf, "%s-p%02d-f64-c%u-v%u.bin", kind_name(kind), p, kConventionVersion,
                kFormatVersion);
  return buf;
}

// The payload of (kind, p) from `dir`, or false (missing or rejected file).
bool load(const std::string &dir, int kind, int p, std::vector<double> &out) {
  const std::string path = dir + "/" + file_name(kind, p);
  FILE *f = std::fopen(path.c_str(), "rb");
  if (!f) return false;
  std::vector<unsigned char> bytes;
  unsigned char chunk[1 << 16];
  for (size_t got; (got = std::fread(chunk, 1, sizeof chunk, f)) > 0;) bytes.insert(bytes.end(), chunk, chunk + got);
  std::fclose(f);
  if (bytes.size() < kHeaderLen) return false;
  const unsigned char *h = bytes.data();
  if (std::memcmp(h, kMagic, 8) != 0) return false;
  if (get32(h + 8) != kFormatVersion || get32(h + 12) != kConventionVersion) return false;
  if (get32(h + 16) != uint32_t(kind) || get32(h + 20) != uint32_t(p) || get32(h + 24) != 64) return false;
  const uint64_t len = get64(h + 28);
  const size_t n2 = size_t(ncoef(p)) * ncoef(p), count = (kind == 3 ? 316 : 8) * n2;
  if (len != bytes.size() - kHeaderLen || len != count * sizeof(double)) return false;
  if (fnv1a(h + kHeaderLen, len) != get64(h + 36)) return false;
  out.resize(count);
  // little-endian IEEE 754 values (this path runs on little-endian hosts only)
  std::memcpy(out.data(), h + kHeaderLen, len);
  return true;
}

void store(const std::string &dir, int kind, int p, const std::vector<double> &v) {
  const std::string name = file_name(kind, p);
  const std::string tmp = dir + "/." + name + "." + std::to_string(getpid()) + "-" +
                          std::to_string(g_counter.fetch_add(1)) + ".tmp";
  const size_t len = v.size() * sizeof(double);
  std::vector<unsigned char> bytes(kHeaderLen + len);
  std::memcpy(bytes.data() + kHeaderLen, v.data(), len);
  unsigned char *h = bytes.data();
  std::memcpy(h, kMagic, 8);
  put32(h + 8, kFormatVersion);
  put32(h + 12, kConventionVersion);
  put32(h + 16, uint32_t(kind));
  put32(h + 20, uint32_t(p));
  put32(h + 24, 64);
  put64(h + 28, len);
  put64(h + 36, fnv1a(h + kHeaderLen, len));
  const int fd = ::open(tmp.c_str(), O_WRONLY | O_CREAT | O_EXCL, 0644);
  if (fd < 0) return;  // an unwritable directory: the tables are still built
  size_t done = 0;
  while (done < bytes.size()) {
    const ssize_t w = ::write(fd, bytes.data() + done, bytes.size() - done);
    if (w <= 0) break;
    done += size_t(w);
  }
  const bool ok = done == bytes.size() && ::fsync(fd) == 0;
  ::close(fd);
  if (!ok || std::rename(tmp.c_str(), (dir + "/" + name).c_str()) != 0) std::remove(tmp.c_str());
}

}  // namespace

void set_table_cache(const char *directory) {
  std::lock_guard<std::mutex> lock(g_mu);
  g_dir = directory ? directory : "";
  g_hits = g_misses = 0;
}

void table_cache_stats(int64_t *hits, int64_t *misses) {
  std::lock_guard<std::mutex> lock(g_mu);
  if (hits) *hits = g_hits;
  if (misses) *misses = g_misses;
}

void cached_tables(int p, std::vector<double> &m2m, std::vector<double> &l2l, std::vector<double> &m2l) {
  std::string dir;
  {
    std::lock_guard<std::mutex> lock(g_mu);
    dir = g_dir;
  }
  if (dir.empty()) {
    build_tables(p, m2m, l2l, m2l);
    return;
  }
  const bool hit = load(dir, 1, p, m2m) && load(dir, 2, p, l2l) && load(dir, 3, p, m2l);
  {
    std::lock_guard<std::mutex> lock(g_mu);
    ++(hit ? g_hits : g_misses);
  }
  if (hit) return;
  build_tables(p, m2m, l2l, m2l);
  store(dir, 1, p, m2m);
  store(dir, 2, p, l2l);
  store(dir, 3, p, m2l);
}

}  // namespace fmmh
