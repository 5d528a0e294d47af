// World snapshot file format (host side), shared by the snapshot reader/writer and the streaming
// device upload.
//
// The reference keeps worlds in memory only (ARCHITECTURE.md:243 lists disk streaming as future
// work), so this format is new. It stores what the pools are built from, each non-empty chunk's
// local 64-tree exactly as ChunkTree64 holds it (nodes[0] = root, chunk-local pointers, then the
// chunk's bricks), in the Morton order of rebuild_all_gpu_pools (world_map.rs:158-205). The pools
// are a pure function of that sequence, so a loader can place each chunk at its final pool offset
// as it streams and relocate its pointers there.
//
//   offset 0    SnapHeader (64 B)
//   offset 64   SnapChunk[n_present] (32 B each), ascending Morton code of the chunk coordinates
//   payloads    per chunk: n_nodes x 16 B nodes, n_bricks x 64 B bricks, zero-padded to `align`;
//               every payload starts at a multiple of `align` (4096: page-aligned for mmap and
//               direct reads)
//
// All integers little-endian. Checksums are CRC-32 (IEEE 802.3, reflected polynomial 0xEDB88320,
// initial value and final xor 0xFFFFFFFF: the zlib / PNG CRC).
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>

#include "../../include/arcadia_world.h"

namespace arc {

constexpr char kSnapMagic[8] = {'A', 'R', 'C', 'W', '6', '4', 'T', '\0'};
constexpr uint32_t kSnapVersion = 1u;
constexpr uint32_t kSnapAlign = 4096u;

struct SnapHeader {
  char magic[8];
  uint32_t version;
  uint32_t header_bytes;  // 64
  uint32_t grid[3];
  uint32_t n_present;     // chunks stored (non-empty trees)
  uint64_t total_nodes;   // sum of n_nodes over stored chunks
  uint64_t total_bricks;  // sum of n_bricks over stored chunks
  uint32_t align;         // payload alignment
  uint32_t index_crc;     // CRC-32 of the SnapChunk table
  uint32_t reserved;      // 0
  uint32_t header_crc;    // CRC-32 of bytes [0, 60)
};
static_assert(sizeof(SnapHeader) == 64, "snapshot header is 64 bytes");

struct SnapChunk {
  uint32_t chunk;     // flat index x + y*gx + z*gx*gy
  uint32_t n_nodes;   // >= 1; nodes[0] is the root, with a non-empty child mask
  uint32_t n_bricks;
  uint32_t crc;       // CRC-32 of the payload (nodes then bricks, without padding)
  uint64_t offset;    // byte offset of the payload
  uint64_t solid;     // solid voxels of the chunk
};
static_assert(sizeof(SnapChunk) == 32, "snapshot index entry is 32 bytes");

// An open snapshot: the file mapped read-only, header and index validated.
struct Snapshot {
  int fd = -1;
  const uint8_t* base = nullptr;
  size_t size = 0;
  SnapHeader h{};
  const SnapChunk* index = nullptr;
  std::string path;
};

uint32_t crc32(const void* data, size_t n, uint32_t crc = 0u);

// Payload of chunk entry k and its byte size.
inline const uint8_t* snap_payload(const Snapshot& s, uint32_t k) { return s.base + s.index[k].offset; }
inline size_t snap_payload_bytes(const SnapChunk& e) {
  return (size_t)e.n_nodes * sizeof(arc_node64) + (size_t)e.n_bricks * sizeof(arc_brick64);
}

// Checks entry k's payload: CRC, a non-empty root, and the tree rules the pools require (pointers
// inside the chunk's own arrays, at most four levels with leaves at the last). ARC_OK or
// ARC_ERR_CORRUPT with the reason in `why`.
int snap_check_chunk(const Snapshot& s, uint32_t k, std::string& why);

// The tree rules of the upload check, over one tree rooted at `root`.
int check_tree(const arc_node64* nodes, uint64_t n_nodes, uint64_t n_bricks, uint32_t root, std::string& why);

// Whether any node's lod_payload material or any brick byte is `mat`. The traversal reads a
// material from exactly those two places (lod_payload >> 24 of uniform nodes and LOD stops, brick
// bytes of mixed leaves), so pools where this is false for water trace identically with and
// without ignore_water.
bool pools_hold_material(const arc_node64* nodes, uint64_t n_nodes, const arc_brick64* bricks, uint64_t n_bricks,
                         uint32_t mat);

}  // namespace arc
