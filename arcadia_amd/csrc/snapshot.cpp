// World snapshots: writer, validating reader (mmap) and host loader. The streaming device upload
// is arc_snapshot_upload in api.cpp. The file layout is described in snapshot.h.
#include "snapshot.h"

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <cstdio>
#include <cstring>
#include <new>
#include <thread>
#include <vector>

#include "world_internal.h"

namespace arc {

// ---- CRC-32, slicing-by-8 (Kounavis and Berry, "A Systematic Approach to Building High
// Performance Software-Based CRC Generators", 2005): eight 256-entry tables, T[0] the bytewise
// table of the reflected polynomial, T[k][i] = (T[k-1][i] >> 8) ^ T[0][T[k-1][i] & 0xFF].
namespace {
struct CrcTables {
  uint32_t t[8][256];
  CrcTables() {
    for (uint32_t i = 0; i < 256; ++i) {
      uint32_t c = i;
      for (int k = 0; k < 8; ++k) c = (c & 1u) ? (c >> 1) ^ 0xEDB88320u : c >> 1;
      t[0][i] = c;
    }
    for (int k = 1; k < 8; ++k)
      for (uint32_t i = 0; i < 256; ++i) t[k][i] = (t[k - 1][i] >> 8) ^ t[0][t[k - 1][i] & 0xFFu];
  }
};
const CrcTables& crc_tables() {
  static const CrcTables tables;
  return tables;
}
}  // namespace

uint32_t crc32(const void* data, size_t n, uint32_t crc) {
  const CrcTables& T = crc_tables();
  const uint8_t* p = static_cast<const uint8_t*>(data);
  uint32_t c = ~crc;
  while (n >= 8) {
    uint32_t lo, hi;
    std::memcpy(&lo, p, 4);
    std::memcpy(&hi, p + 4, 4);
    lo ^= c;
    c = T.t[7][lo & 0xFFu] ^ T.t[6][(lo >> 8) & 0xFFu] ^ T.t[5][(lo >> 16) & 0xFFu] ^ T.t[4][lo >> 24] ^
        T.t[3][hi & 0xFFu] ^ T.t[2][(hi >> 8) & 0xFFu] ^ T.t[1][(hi >> 16) & 0xFFu] ^ T.t[0][hi >> 24];
    p += 8;
    n -= 8;
  }
  while (n--) c = (c >> 8) ^ T.t[0][(c ^ *p++) & 0xFFu];
  return ~c;
}

bool pools_hold_material(const arc_node64* nodes, uint64_t n_nodes, const arc_brick64* bricks, uint64_t n_bricks,
                         uint32_t mat) {
  for (uint64_t i = 0; i < n_nodes; ++i)
    if ((nodes[i].lod_payload >> 24) == mat) return true;
  const uint8_t* b = reinterpret_cast<const uint8_t*>(bricks);
  return n_bricks != 0 && std::memchr(b, (int)mat, (size_t)n_bricks * sizeof(arc_brick64)) != nullptr;
}

// Structural check of a tree: every reachable node's children lie inside the node array, mixed
// leaves point inside the brick array, and the tree is at most four levels deep with leaves at
// the last level (the invariants the traversal kernels rely on).
int check_tree(const arc_node64* nodes, uint64_t n_nodes, uint64_t n_bricks, uint32_t root, std::string& why) {
  struct Item {
    uint32_t idx, level;
  };
  std::vector<Item> stack;
  stack.push_back({root, 0});
  while (!stack.empty()) {
    Item it = stack.back();
    stack.pop_back();
    if (it.idx >= n_nodes) {
      why = "node index " + std::to_string(it.idx) + " outside the node pool";
      return ARC_ERR_INVALID;
    }
    const arc_node64& n = nodes[it.idx];
    const uint64_t mask = (uint64_t)n.mask_low | ((uint64_t)n.mask_high << 32);
    if (mask == 0) continue;
    const uint32_t base = n.child_ptr & ARC_PTR_INDEX_MASK;
    if (n.child_ptr & ARC_IS_LEAF_FLAG) {
      if (!(n.child_ptr & ARC_IS_UNIFORM_FLAG) && base >= n_bricks) {
        why = "brick index " + std::to_string(base) + " outside the brick pool";
        return ARC_ERR_INVALID;
      }
      continue;
    }
    if (it.level >= 3) {
      why = "internal node " + std::to_string(it.idx) + " below level 3";
      return ARC_ERR_INVALID;
    }
    const uint32_t cnt = (uint32_t)__builtin_popcountll(mask);
    if ((uint64_t)base + cnt > n_nodes) {
      why = "children of node " + std::to_string(it.idx) + " outside the node pool";
      return ARC_ERR_INVALID;
    }
    for (uint32_t k = 0; k < cnt; ++k) stack.push_back({base + k, it.level + 1});
  }
  return ARC_OK;
}

int snap_check_chunk(const Snapshot& s, uint32_t k, std::string& why) {
  const SnapChunk& e = s.index[k];
  const uint8_t* p = snap_payload(s, k);
  if (crc32(p, snap_payload_bytes(e)) != e.crc) {
    why = "chunk " + std::to_string(e.chunk) + ": payload checksum mismatch";
    return ARC_ERR_CORRUPT;
  }
  // the payload is only 16-byte aligned through `align`; copy-free reads are fine on x86-64
  const arc_node64* nodes = reinterpret_cast<const arc_node64*>(p);
  if (nodes[0].mask_low == 0 && nodes[0].mask_high == 0) {
    why = "chunk " + std::to_string(e.chunk) + ": empty root stored";
    return ARC_ERR_CORRUPT;
  }
  std::string w;
  if (check_tree(nodes, e.n_nodes, e.n_bricks, 0u, w) != ARC_OK) {
    why = "chunk " + std::to_string(e.chunk) + ": " + w;
    return ARC_ERR_CORRUPT;
  }
  return ARC_OK;
}

namespace {

uint32_t chunk_morton(const uint32_t grid[3], uint32_t idx) {
  const uint32_t cx = idx % grid[0], cy = (idx / grid[0]) % grid[1], cz = idx / (grid[0] * grid[1]);
  return morton3d(cx, cy, cz);
}

int validate(Snapshot& s, std::string& why) {
  if (s.size < sizeof(SnapHeader)) {
    why = "file shorter than the header";
    return ARC_ERR_CORRUPT;
  }
  std::memcpy(&s.h, s.base, sizeof s.h);
  const SnapHeader& h = s.h;
  if (std::memcmp(h.magic, kSnapMagic, 8) != 0) {
    why = "not a world snapshot (bad magic)";
    return ARC_ERR_CORRUPT;
  }
  if (h.version != kSnapVersion || h.header_bytes != sizeof(SnapHeader)) {
    why = "unsupported snapshot version " + std::to_string(h.version);
    return ARC_ERR_CORRUPT;
  }
  if (crc32(s.base, offsetof(SnapHeader, header_crc)) != h.header_crc) {
    why = "header checksum mismatch";
    return ARC_ERR_CORRUPT;
  }
  for (int a = 0; a < 3; ++a)
    if (h.grid[a] == 0 || h.grid[a] > 256) {
      why = "grid dims out of range";
      return ARC_ERR_CORRUPT;
    }
  const uint64_t nchunks = (uint64_t)h.grid[0] * h.grid[1] * h.grid[2];
  if (h.align < 64 || h.align > (1u << 20) || (h.align & (h.align - 1)) || h.reserved != 0 ||
      h.n_present > nchunks) {
    why = "bad header fields";
    return ARC_ERR_CORRUPT;
  }
  const uint64_t index_end = sizeof(SnapHeader) + (uint64_t)h.n_present * sizeof(SnapChunk);
  if (index_end > s.size) {
    why = "index runs past the end of the file";
    return ARC_ERR_CORRUPT;
  }
  s.index = reinterpret_cast<const SnapChunk*>(s.base + sizeof(SnapHeader));
  if (crc32(s.index, (size_t)h.n_present * sizeof(SnapChunk)) != h.index_crc) {
    why = "index checksum mismatch";
    return ARC_ERR_CORRUPT;
  }
  uint64_t tn = 0, tb = 0, prev_end = index_end;
  int64_t prev_code = -1;
  for (uint32_t k = 0; k < h.n_present; ++k) {
    const SnapChunk& e = s.index[k];
    if (e.chunk >= nchunks || e.n_nodes == 0) {
      why = "index entry " + std::to_string(k) + ": bad chunk or node count";
      return ARC_ERR_CORRUPT;
    }
    const int64_t code = chunk_morton(h.grid, e.chunk);
    if (code <= prev_code) {
      why = "index entry " + std::to_string(k) + ": chunks not in ascending Morton order";
      return ARC_ERR_CORRUPT;
    }
    prev_code = code;
    const uint64_t bytes = snap_payload_bytes(e);
    // offset and size are tested apart: a forged offset near 2^64 must not wrap the sum
    if (e.offset % h.align != 0 || e.offset < prev_end || e.offset > s.size || bytes > s.size - e.offset) {
      why = "index entry " + std::to_string(k) + ": payload offset out of place";
      return ARC_ERR_CORRUPT;
    }
    prev_end = e.offset + bytes;
    tn += e.n_nodes;
    tb += e.n_bricks;
  }
  if (tn != h.total_nodes || tb != h.total_bricks) {
    why = "node/brick totals disagree with the index";
    return ARC_ERR_CORRUPT;
  }
  if (tn + 1 > (uint64_t)ARC_PTR_INDEX_MASK || tb + 1 > (uint64_t)ARC_PTR_INDEX_MASK) {
    why = "pools larger than the 30-bit pointer range";
    return ARC_ERR_CORRUPT;
  }
  return ARC_OK;
}

void close_snapshot(Snapshot* s) {
  if (!s) return;
  if (s->base) munmap(const_cast<uint8_t*>(s->base), s->size);
  if (s->fd >= 0) close(s->fd);
  delete s;
}

// Runs f(k) for k in [0, n) on up to `nthreads` threads; returns the first non-OK status.
template <class F>
int parallel_chunks(uint32_t n, int nthreads, F f) {
  const uint32_t nt = (uint32_t)std::max(1, std::min<int>(nthreads, (int)std::max(1u, n)));
  std::atomic<uint32_t> next{0};
  std::atomic<int> status{ARC_OK};
  auto work = [&]() {
    for (;;) {
      const uint32_t k = next.fetch_add(1);
      if (k >= n || status.load() != ARC_OK) return;
      const int rc = f(k);
      if (rc != ARC_OK) {
        int ok = ARC_OK;
        status.compare_exchange_strong(ok, rc);
      }
    }
  };
  std::vector<std::thread> th;
  for (uint32_t t = 1; t < nt; ++t) th.emplace_back(work);
  work();
  for (auto& t : th) t.join();
  return status.load();
}

thread_local std::string g_err;

bool write_all(int fd, const void* p, size_t n, uint64_t off) {
  const uint8_t* b = static_cast<const uint8_t*>(p);
  while (n) {
    const ssize_t w = pwrite(fd, b, std::min<size_t>(n, (size_t)1 << 30), (off_t)off);
    if (w < 0) {
      if (errno == EINTR) continue;
      return false;
    }
    b += w;
    off += (uint64_t)w;
    n -= (size_t)w;
  }
  return true;
}

}  // namespace
}  // namespace arc

using namespace arc;

extern "C" {

const char* arc_snapshot_last_error(void) { return g_err.c_str(); }

uint32_t arc_crc32(const void* data, uint64_t n, uint32_t crc) { return data || !n ? crc32(data, n, crc) : 0u; }

int arc_snapshot_write(const arc_world_host* wh, const char* path) {
  const World* w = reinterpret_cast<const World*>(wh);
  if (!w || !path) return ARC_ERR_INVALID;
  const uint32_t nchunks = w->grid[0] * w->grid[1] * w->grid[2];
  // the chunks rebuild_all_gpu_pools places, in its order
  std::vector<std::pair<uint32_t, uint32_t>> order;
  for (uint32_t i = 0; i < nchunks; ++i) {
    const ChunkTree& c = w->chunks[i];
    if (c.nodes.empty() || (c.nodes[0].mask_low == 0 && c.nodes[0].mask_high == 0)) continue;
    order.push_back({chunk_morton(w->grid, i), i});
  }
  std::sort(order.begin(), order.end());
  SnapHeader h{};
  std::memcpy(h.magic, kSnapMagic, 8);
  h.version = kSnapVersion;
  h.header_bytes = sizeof(SnapHeader);
  std::memcpy(h.grid, w->grid, sizeof h.grid);
  h.n_present = (uint32_t)order.size();
  h.align = kSnapAlign;
  std::vector<SnapChunk> index(order.size());
  uint64_t off = sizeof(SnapHeader) + (uint64_t)order.size() * sizeof(SnapChunk);
  for (size_t k = 0; k < order.size(); ++k) {
    const ChunkTree& c = w->chunks[order[k].second];
    SnapChunk& e = index[k];
    e.chunk = order[k].second;
    e.n_nodes = (uint32_t)c.nodes.size();
    e.n_bricks = (uint32_t)c.bricks.size();
    uint32_t crc = crc32(c.nodes.data(), c.nodes.size() * sizeof(arc_node64));
    e.crc = crc32(c.bricks.data(), c.bricks.size() * sizeof(arc_brick64), crc);
    off = (off + kSnapAlign - 1) / kSnapAlign * kSnapAlign;
    e.offset = off;
    e.solid = c.solid;
    off += snap_payload_bytes(e);
    h.total_nodes += e.n_nodes;
    h.total_bricks += e.n_bricks;
  }
  if (h.total_nodes + 1 > (uint64_t)ARC_PTR_INDEX_MASK || h.total_bricks + 1 > (uint64_t)ARC_PTR_INDEX_MASK) {
    g_err = "world larger than the 30-bit pointer range";
    return ARC_ERR_INVALID;
  }
  h.index_crc = crc32(index.data(), index.size() * sizeof(SnapChunk));
  h.header_crc = crc32(&h, offsetof(SnapHeader, header_crc));
  // written to a temporary file, then renamed over `path`: a reader never sees a partial file
  const std::string tmp = std::string(path) + ".tmp";
  const int fd = open(tmp.c_str(), O_WRONLY | O_CREAT | O_TRUNC, 0644);
  if (fd < 0) {
    g_err = "cannot create " + tmp + ": " + std::strerror(errno);
    return ARC_ERR_IO;
  }
  bool ok = write_all(fd, &h, sizeof h, 0) && write_all(fd, index.data(), index.size() * sizeof(SnapChunk), sizeof h);
  for (size_t k = 0; ok && k < order.size(); ++k) {
    const ChunkTree& c = w->chunks[order[k].second];
    const uint64_t nb = c.nodes.size() * sizeof(arc_node64);
    ok = write_all(fd, c.nodes.data(), nb, index[k].offset) &&
         write_all(fd, c.bricks.data(), c.bricks.size() * sizeof(arc_brick64), index[k].offset + nb);
  }
  const uint64_t end = (off + kSnapAlign - 1) / kSnapAlign * kSnapAlign;
  ok = ok && ftruncate(fd, (off_t)end) == 0 && fsync(fd) == 0;
  ok = (close(fd) == 0) && ok;
  if (!ok || rename(tmp.c_str(), path) != 0) {
    g_err = "writing " + std::string(path) + " failed: " + std::strerror(errno);
    unlink(tmp.c_str());
    return ARC_ERR_IO;
  }
  return ARC_OK;
}

int arc_snapshot_open(const char* path, arc_snapshot** out) {
  if (!path || !out) return ARC_ERR_INVALID;
  *out = nullptr;
  Snapshot* s = new (std::nothrow) Snapshot();
  if (!s) return ARC_ERR_INVALID;
  s->path = path;
  s->fd = open(path, O_RDONLY);
  struct stat st;
  if (s->fd < 0 || fstat(s->fd, &st) != 0) {
    g_err = "cannot open " + std::string(path) + ": " + std::strerror(errno);
    close_snapshot(s);
    return ARC_ERR_IO;
  }
  s->size = (size_t)st.st_size;
  if (s->size == 0) {
    g_err = std::string(path) + ": empty file";
    close_snapshot(s);
    return ARC_ERR_CORRUPT;
  }
  void* m = mmap(nullptr, s->size, PROT_READ, MAP_PRIVATE, s->fd, 0);
  if (m == MAP_FAILED) {
    g_err = "cannot map " + std::string(path) + ": " + std::strerror(errno);
    close_snapshot(s);
    return ARC_ERR_IO;
  }
  s->base = static_cast<const uint8_t*>(m);
  madvise(m, s->size, MADV_SEQUENTIAL);
  std::string why;
  const int rc = validate(*s, why);
  if (rc != ARC_OK) {
    g_err = std::string(path) + ": " + why;
    close_snapshot(s);
    return rc;
  }
  *out = reinterpret_cast<arc_snapshot*>(s);
  return ARC_OK;
}

void arc_snapshot_close(arc_snapshot* s) { close_snapshot(reinterpret_cast<Snapshot*>(s)); }

int arc_snapshot_info(const arc_snapshot* sh, uint32_t grid[3], uint32_t* n_present, uint64_t* total_nodes,
                      uint64_t* total_bricks) {
  const Snapshot* s = reinterpret_cast<const Snapshot*>(sh);
  if (!s) return ARC_ERR_INVALID;
  if (grid) std::memcpy(grid, s->h.grid, sizeof s->h.grid);
  if (n_present) *n_present = s->h.n_present;
  if (total_nodes) *total_nodes = s->h.total_nodes;
  if (total_bricks) *total_bricks = s->h.total_bricks;
  return ARC_OK;
}

int arc_snapshot_verify(const arc_snapshot* sh, int nthreads) {
  const Snapshot* s = reinterpret_cast<const Snapshot*>(sh);
  if (!s) return ARC_ERR_INVALID;
  std::vector<std::string> why(s->h.n_present);
  const int rc = parallel_chunks(s->h.n_present, nthreads, [&](uint32_t k) { return snap_check_chunk(*s, k, why[k]); });
  if (rc != ARC_OK)
    for (auto& w : why)
      if (!w.empty()) {
        g_err = s->path + ": " + w;
        break;
      }
  return rc;
}

int arc_snapshot_load_host(const arc_snapshot* sh, int nthreads, arc_world_host** out) {
  const Snapshot* s = reinterpret_cast<const Snapshot*>(sh);
  if (!s || !out) return ARC_ERR_INVALID;
  *out = nullptr;
  arc_world_host* wh = nullptr;
  int rc = arc_world_host_create(s->h.grid, &wh);
  if (rc != ARC_OK) return rc;
  World* w = reinterpret_cast<World*>(wh);
  std::vector<std::string> why(s->h.n_present);
  rc = parallel_chunks(s->h.n_present, nthreads, [&](uint32_t k) {
    const int r = snap_check_chunk(*s, k, why[k]);
    if (r != ARC_OK) return r;
    const SnapChunk& e = s->index[k];
    ChunkTree& c = w->chunks[e.chunk];
    const uint8_t* p = snap_payload(*s, k);
    c.nodes.resize(e.n_nodes);
    c.bricks.resize(e.n_bricks);
    std::memcpy(c.nodes.data(), p, (size_t)e.n_nodes * sizeof(arc_node64));
    std::memcpy(c.bricks.data(), p + (size_t)e.n_nodes * sizeof(arc_node64), (size_t)e.n_bricks * sizeof(arc_brick64));
    c.solid = e.solid;
    return ARC_OK;
  });
  if (rc != ARC_OK) {
    for (auto& m : why)
      if (!m.empty()) {
        g_err = s->path + ": " + m;
        break;
      }
    arc_world_host_destroy(wh);
    return rc;
  }
  *out = wh;
  return ARC_OK;
}

}  // extern "C"
