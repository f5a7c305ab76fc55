// Binary CSR cache: one data graph as an mmap-able file, so a large graph
// is loaded in the time it takes to page it in instead of being parsed from
// the text format (io/IO.cpp:46-100 reads it token by token with fscanf;
// at 64M edges that is minutes).
//
// Layout, little-endian:
//   [0, 8)    magic "GPSMCSR" + format version byte 1
//   [8, 12)   n (u32)
//   [12, 16)  zero
//   [16, 24)  adj = row_offset[n] (u64)
//   [24, 64)  zero
//   [64, ...) row_offset[n+1] u32, col[adj] u32, labels[n] u32
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>

#include "gpsm.h"
#include "gpsm_gen.h"

namespace {

const char kMagic[8] = {'G', 'P', 'S', 'M', 'C', 'S', 'R', 1};
constexpr size_t kHeader = 64;

std::mutex g_maps_mu;
std::map<const void*, std::pair<void*, size_t>> g_maps;  // row_offset -> (mapping, length)

size_t file_bytes(uint32_t n, uint64_t adj) { return kHeader + 4 * ((size_t)n + 1 + adj + n); }

}  // namespace

extern "C" int gpsm_io_write_csr(const char* path, uint32_t n, const uint32_t* row_offset, const uint32_t* col,
                                 const uint32_t* labels) {
  if (!path || !row_offset || !labels || (row_offset[n] && !col)) return GPSM_EINVAL;
  unsigned char h[kHeader] = {0};
  std::memcpy(h, kMagic, 8);
  std::memcpy(h + 8, &n, 4);
  const uint64_t adj = row_offset[n];
  std::memcpy(h + 16, &adj, 8);
  FILE* f = std::fopen(path, "wb");
  if (!f) return GPSM_EINVAL;
  bool ok = std::fwrite(h, 1, kHeader, f) == kHeader &&
            std::fwrite(row_offset, 4, (size_t)n + 1, f) == (size_t)n + 1 &&
            (adj == 0 || std::fwrite(col, 4, adj, f) == adj) && std::fwrite(labels, 4, n, f) == n;
  ok = (std::fclose(f) == 0) && ok;
  return ok ? GPSM_OK : GPSM_EINVAL;
}

extern "C" int gpsm_io_map_csr(const char* path, gpsm_csr* out) {
  if (!path || !out) return GPSM_EINVAL;
  std::memset(out, 0, sizeof(*out));
  const int fd = open(path, O_RDONLY);
  if (fd < 0) return GPSM_EINVAL;
  struct stat st;
  if (fstat(fd, &st) != 0 || (size_t)st.st_size < kHeader) {
    close(fd);
    return GPSM_EFORMAT;
  }
  void* base = mmap(nullptr, (size_t)st.st_size, PROT_READ, MAP_PRIVATE, fd, 0);
  close(fd);
  if (base == MAP_FAILED) return GPSM_EINVAL;
  const unsigned char* h = static_cast<const unsigned char*>(base);
  uint32_t n;
  uint64_t adj;
  std::memcpy(&n, h + 8, 4);
  std::memcpy(&adj, h + 16, 8);
  const size_t len = (size_t)st.st_size;
  bool ok = std::memcmp(h, kMagic, 8) == 0 && n > 0 && adj < (1ull << 32) && file_bytes(n, adj) == len;
  const uint32_t* ro = reinterpret_cast<const uint32_t*>(h + kHeader);
  if (ok) {
    // offsets must be monotone and end at adj, so that every later read of
    // a row stays inside the file
    ok = ro[0] == 0 && ro[n] == adj;
    for (uint32_t v = 0; ok && v < n; ++v) ok = ro[v] <= ro[v + 1];
  }
  if (!ok) {
    munmap(base, len);
    return GPSM_EFORMAT;
  }
  madvise(base, len, MADV_SEQUENTIAL);
  out->n = n;
  out->adj = adj;
  out->row_offset = const_cast<uint32_t*>(ro);
  out->col = const_cast<uint32_t*>(ro + n + 1);
  out->labels = const_cast<uint32_t*>(ro + n + 1 + adj);
  std::lock_guard<std::mutex> lock(g_maps_mu);
  g_maps[out->row_offset] = {base, len};
  return GPSM_OK;
}

extern "C" void gpsm_io_unmap_csr(gpsm_csr* g) {
  if (!g || !g->row_offset) return;
  std::lock_guard<std::mutex> lock(g_maps_mu);
  auto it = g_maps.find(g->row_offset);
  if (it == g_maps.end()) return;  // not a mapping of this library
  munmap(it->second.first, it->second.second);
  g_maps.erase(it);
  std::memset(g, 0, sizeof(*g));
}
