// Patch-batch planning for arc_world_patch and its host model arc_world_patch_host.
//
// Renderer::apply_world_patches (crates/arcadia-render/src/renderer.rs:147-315) writes each
// GpuPatch's root, node slice and brick slice in queue order with queue.write_buffer; a later
// patch overwrites what an earlier one wrote. Here the batch becomes one staging image and one
// scatter launch: walking the patches from last to first, each keeps only the parts of its slices
// that no later patch covers, so the pieces are disjoint and their order no longer matters.
#include "patch_plan.h"

#include <algorithm>
#include <cstdio>
#include <cstring>

namespace arcplan {
namespace {

struct Span {
  uint64_t lo, hi;  // [lo, hi) in pool elements
};

// Appends to `out` the parts of [lo, hi) not in `covered` (sorted, disjoint), then adds [lo, hi).
void uncovered_parts(std::vector<Span>& covered, uint64_t lo, uint64_t hi, std::vector<Span>& out) {
  uint64_t cur = lo;
  auto it = std::lower_bound(covered.begin(), covered.end(), lo, [](const Span& s, uint64_t v) { return s.hi <= v; });
  for (auto j = it; j != covered.end() && j->lo < hi; ++j) {
    if (j->lo > cur) out.push_back({cur, j->lo});
    cur = std::max(cur, j->hi);
  }
  if (cur < hi) out.push_back({cur, hi});
  Span m{lo, hi};
  auto first = std::lower_bound(covered.begin(), covered.end(), lo, [](const Span& s, uint64_t v) { return s.hi < v; });
  auto last = first;
  while (last != covered.end() && last->lo <= hi) {
    m.lo = std::min(m.lo, last->lo);
    m.hi = std::max(m.hi, last->hi);
    ++last;
  }
  first = covered.erase(first, last);
  covered.insert(first, m);
}

std::string fmt(const char* f, unsigned long long a, unsigned long long b = 0) {
  char buf[256];
  std::snprintf(buf, sizeof buf, f, a, b);
  return buf;
}

}  // namespace

int plan_patches(const arc_gpu_patch* patches, uint64_t n, uint64_t nchunks, uint64_t node_cap, uint64_t brick_cap,
                 Plan& plan, std::string& err) {
  plan = Plan{};
  uint64_t need_nodes = node_cap, need_bricks = brick_cap;
  for (uint64_t i = 0; i < n; ++i) {
    const arc_gpu_patch& p = patches[i];
    if (p.chunk_flat_idx >= nchunks && p.chunk_flat_idx != ARC_EMPTY_CHUNK_SENTINEL) {
      err = fmt("patch %llu: chunk index out of range", i);
      return ARC_ERR_INVALID;
    }
    if ((p.n_nodes && !p.nodes) || (p.n_bricks && !p.bricks)) {
      err = fmt("patch %llu: null slice", i);
      return ARC_ERR_INVALID;
    }
    const uint64_t ne = (uint64_t)p.node_offset + p.n_nodes, be = (uint64_t)p.brick_offset + p.n_bricks;
    if (ne > ARC_PTR_INDEX_MASK || be > ARC_PTR_INDEX_MASK) {
      err = fmt("patch %llu beyond the 30-bit pointer range", i);
      return ARC_ERR_CAPACITY;
    }
    need_nodes = std::max(need_nodes, ne);
    need_bricks = std::max(need_bricks, be);
  }
  // Every pointer a patch writes stays inside the grown pools, so no traversal can read outside
  // them whatever the patch contents.
  for (uint64_t i = 0; i < n; ++i) {
    const arc_gpu_patch& p = patches[i];
    if (p.chunk_flat_idx != ARC_EMPTY_CHUNK_SENTINEL && p.new_root_node_idx != ARC_EMPTY_CHUNK_SENTINEL &&
        p.new_root_node_idx >= need_nodes) {
      err = fmt("patch %llu: root %llu outside the node pool", i, p.new_root_node_idx);
      return ARC_ERR_INVALID;
    }
    for (uint32_t k = 0; k < p.n_nodes; ++k) {
      const arc_node64& nd = p.nodes[k];
      const uint64_t mask = (uint64_t)nd.mask_low | ((uint64_t)nd.mask_high << 32);
      if (mask == 0) continue;
      const uint64_t base = nd.child_ptr & ARC_PTR_INDEX_MASK;
      if (nd.child_ptr & ARC_IS_LEAF_FLAG) {
        if (!(nd.child_ptr & ARC_IS_UNIFORM_FLAG) && base >= need_bricks) {
          err = fmt("patch %llu node %llu: brick outside the brick pool", i, k);
          return ARC_ERR_INVALID;
        }
      } else if (base + (uint64_t)__builtin_popcountll(mask) > need_nodes) {
        err = fmt("patch %llu node %llu: children outside the node pool", i, k);
        return ARC_ERR_INVALID;
      }
    }
  }
  plan.need_nodes = need_nodes;
  plan.need_bricks = need_bricks;
  std::vector<Span> cov_n, cov_b, tmp;
  std::vector<uint8_t> root_seen(nchunks, 0);
  for (uint64_t ii = n; ii-- > 0;) {
    const arc_gpu_patch& p = patches[ii];
    if (p.chunk_flat_idx != ARC_EMPTY_CHUNK_SENTINEL && !root_seen[p.chunk_flat_idx]) {
      root_seen[p.chunk_flat_idx] = 1;
      plan.roots.push_back({p.chunk_flat_idx, p.new_root_node_idx});
    }
    if (p.n_nodes) {
      tmp.clear();
      uncovered_parts(cov_n, p.node_offset, (uint64_t)p.node_offset + p.n_nodes, tmp);
      for (const Span& s : tmp)
        plan.parts.push_back({p.nodes + (s.lo - p.node_offset), s.lo, s.hi - s.lo, 0});
    }
    if (p.n_bricks) {
      tmp.clear();
      uncovered_parts(cov_b, p.brick_offset, (uint64_t)p.brick_offset + p.n_bricks, tmp);
      for (const Span& s : tmp)
        plan.parts.push_back({p.bricks + (s.lo - p.brick_offset), s.lo, s.hi - s.lo, 1});
    }
  }
  uint64_t data_units = 0;
  for (const Plan::Src& s : plan.parts) {
    const uint64_t upe = s.pool == 0 ? 1 : 4;  // 16 B units per element
    const uint64_t units = s.n * upe;
    for (uint64_t off = 0; off < units; off += kPieceUnits) {
      Piece pc;
      pc.dst_unit = s.dst * upe + off;
      pc.src_unit = (uint32_t)(data_units + off);  // relative to the data block; rebased below
      pc.n_units = (uint16_t)std::min<uint64_t>(kPieceUnits, units - off);
      pc.pool = s.pool;
      plan.pieces.push_back(pc);
    }
    data_units += units;
    if (s.pool == 0) plan.max_node_written = std::max(plan.max_node_written, s.dst + s.n);
    else plan.max_brick_written = std::max(plan.max_brick_written, s.dst + s.n);
  }
  plan.header_units = plan.pieces.size() + (plan.roots.size() + 1) / 2;
  plan.total_units = plan.header_units + data_units;
  if (plan.total_units > 0xFFFFFFFFull) {
    err = "patch batch larger than 64 GiB";
    return ARC_ERR_CAPACITY;
  }
  for (Piece& pc : plan.pieces) pc.src_unit += (uint32_t)plan.header_units;
  return ARC_OK;
}

void fill_staging(const Plan& plan, void* staging) {
  uint8_t* h = static_cast<uint8_t*>(staging);
  if (!plan.pieces.empty()) std::memcpy(h, plan.pieces.data(), plan.pieces.size() * sizeof(Piece));
  if (!plan.roots.empty())
    std::memcpy(h + plan.pieces.size() * 16, plan.roots.data(), plan.roots.size() * sizeof(RootWrite));
  uint8_t* d = h + plan.header_units * 16;
  for (const Plan::Src& s : plan.parts) {
    const size_t nb = s.n * (s.pool == 0 ? sizeof(arc_node64) : sizeof(arc_brick64));
    std::memcpy(d, s.p, nb);
    d += nb;
  }
}

void scatter_host(const void* staging, uint32_t n_pieces, uint32_t n_roots, uint32_t* roots, void* nodes,
                  void* bricks) {
  const uint8_t* h = static_cast<const uint8_t*>(staging);
  const Piece* pieces = reinterpret_cast<const Piece*>(h);
  for (uint32_t b = 0; b < n_pieces; ++b) {
    const Piece& p = pieces[b];
    uint8_t* dst = static_cast<uint8_t*>(p.pool == 0 ? nodes : bricks) + p.dst_unit * 16;
    std::memcpy(dst, h + (uint64_t)p.src_unit * 16, (size_t)p.n_units * 16);
  }
  const RootWrite* rw = reinterpret_cast<const RootWrite*>(h + (uint64_t)n_pieces * 16);
  for (uint32_t i = 0; i < n_roots; ++i) roots[rw[i].chunk] = rw[i].root;
}

}  // namespace arcplan

extern "C" int arc_world_patch_host(const arc_gpu_patch* patches, uint64_t n, uint32_t n_chunks, uint32_t* roots,
                                    arc_node64* nodes, uint64_t node_cap, arc_brick64* bricks, uint64_t brick_cap) {
  if (n == 0) return ARC_OK;
  if (!patches || !roots || !nodes || !bricks) return ARC_ERR_INVALID;
  arcplan::Plan plan;
  std::string err;
  const int rc = arcplan::plan_patches(patches, n, n_chunks, node_cap, brick_cap, plan, err);
  if (rc != ARC_OK) return rc;
  if (plan.need_nodes > node_cap || plan.need_bricks > brick_cap) return ARC_ERR_CAPACITY;
  std::vector<uint8_t> staging(plan.total_units * 16);
  arcplan::fill_staging(plan, staging.data());
  arcplan::scatter_host(staging.data(), (uint32_t)plan.pieces.size(), (uint32_t)plan.roots.size(), roots, nodes,
                        bricks);
  return ARC_OK;
}
