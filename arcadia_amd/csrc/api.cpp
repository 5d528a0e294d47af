// C ABI of the traversal library: context, device world pools, trace entry points.
// Host code only (built with hipcc as HIP so it can call the launchers).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <initializer_list>
#include <new>
#include <string>
#include <thread>
#include <utility>
#include <vector>

#include "../../include/arcadia_hip.h"
#include "../../include/arcadia_world.h"
#include "gbuffer.h"
#include "patch.h"
#include "patch_plan.h"
#include "pathtrace.h"
#include "restir.h"
#include "snapshot.h"
#include "trace_launch.h"

struct arc_ctx {
  int device = 0;
  uint32_t n_cu = 0;
  uint32_t lds_per_block = 0;
  std::string err;
  // device pools (the @group(0) bindings 1..3)
  uint32_t grid[3] = {0, 0, 0};
  uint32_t* d_roots = nullptr;
  arc_node64* d_nodes = nullptr;
  arc_brick64* d_bricks = nullptr;
  uint64_t n_nodes = 0, node_cap = 0, n_bricks = 0, brick_cap = 0;
  bool has_world = false;
  // The pools may hold water (ARC_MAT_WATER as a node's lod material or a brick byte). When they
  // cannot, ignore_water changes no result and traces run the kernels without the water tests.
  bool water = true;
  // persistent queues (kQueueRegionWords words each): the traversal kernel's, the path kernel's,
  // the shadow walkers', then two for arc_trace_bounce_device
  uint32_t* d_queue = nullptr;
  // scratch for the host-pointer entry point
  arc_ray* d_rays = nullptr;
  arc_hit* d_hits = nullptr;
  uint64_t scratch_n = 0;
  arc_launch_config cfg[3]{};  // per ray class (lod_bias 0, 1, >=2)
  arc_launch_config cfg_bounce{};  // arc_trace_bounce_device
  arc_launch_config cfg_shadow{};  // arc_trace_shadow_device
  // arc_trace_bounce_device: one message word per primary ray (zero between launches)
  uint64_t* d_msg = nullptr;
  uint64_t msg_n = 0;
  // primary G-buffer pass: material palette and per-pass scratch
  float* d_palette = nullptr;
  arc_ray* d_gb_rays = nullptr;
  arc_trace_result* d_gb_res = nullptr;
  uint64_t gb_n = 0;
  // rigid-body instances of the last arc_trace_rigid_device call (ARC_MAX_RIGID_BODIES slots)
  arc_rigid_body* d_bodies = nullptr;
  // bumped whenever the device pool addresses change (upload, pool growth): captured graphs bake them
  uint64_t world_gen = 0;
  // arc_world_patch staging (pinned host + device copy of the same size)
  void* h_stage = nullptr;
  void* d_stage = nullptr;
  size_t stage_bytes = 0;
  // arc_snapshot_upload staging: two pinned batch buffers and their device chunk tables, kept
  // for later uploads (page-locking is a large part of a cold upload)
  uint8_t* h_snap[2] = {nullptr, nullptr};
  uint4* d_snap_table[2] = {nullptr, nullptr};
  size_t snap_bytes = 0, snap_table_bytes = 0;
  // ReSTIR pass scratch: per-pixel ray and result buffers (grown to the largest frame seen)
  uint8_t* d_restir = nullptr;
  size_t restir_bytes = 0;
};

namespace {

int fail(arc_ctx* c, int code, const char* fmt, ...) {
  if (c) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    c->err = buf;
  }
  return code;
}

#define HIP_TRY(ctx, expr)                                                                   \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess) return fail((ctx), ARC_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

constexpr size_t kQueueBytes = 5u * arcdev::kQueueRegionWords * sizeof(uint32_t);

// Traversal kernel, measured on MI355X (full-frame world, ms per primary / bounce launch, refill
// at 16 idle lanes, windows of 64 rays): 6 workgroups of 256 per CU 0.450 / 0.429; 5: 0.464 /
// 0.437; 7: 0.451 / 0.436. Windows of 32 at 6 workgroups: 0.452 / 0.459.
// blocks_per_cu 0 = as many workgroups as the kernel's occupancy keeps resident (the path kernel's
// default: its register budget allows fewer).
arc_launch_config default_config() {
  arc_launch_config c;
  c.blocks_per_cu = 6;
  c.block_threads = 256;
  c.refill_lanes = 16;
  c.batch_rays = 64;
  return c;
}

uint32_t lod_class(const arc_trace_params* p) { return p->lod_bias == 0u ? 0u : (p->lod_bias == 1u ? 1u : 2u); }

bool valid_params(const arc_trace_params* p) {
  if (!p) return false;
  if (p->semantics != ARC_SEM_WORLD && p->semantics != ARC_SEM_CHUNK) return false;
  return true;
}

}  // namespace

// Times every candidate launch shape with `run` (one launch on `stream`, reading the shape from
// *slot), keeps the fastest in *slot and returns it with its median time. Candidates: workgroup
// size, resident workgroups per CU (0 = occupancy limit), refill threshold (idle lanes before a
// wave refills) and window size per queue claim.
namespace {

// Launch-shape autotuner shared by the traversal classes, the one-launch frame and the batched
// views: `run` launches with the shape in *slot; threads x blocks x refill x batch candidates,
// `skip(threads, blocks)` leaves out shapes that cannot launch.
template <typename Run, typename Skip>
int tune_shapes(arc_ctx* c, arc_launch_config* slot, void* stream, uint32_t reps, Run run,
                arc_launch_config* best, float* best_ms, std::initializer_list<uint32_t> kThreads,
                std::initializer_list<uint32_t> kBlocks, Skip skip) {
  if (reps == 0) reps = 3;
  if (reps > 64) reps = 64;
  HIP_TRY(c, hipSetDevice(c->device));
  const arc_launch_config saved = *slot;
  hipStream_t s = (hipStream_t)stream;
  hipEvent_t e0, e1;
  HIP_TRY(c, hipEventCreate(&e0));
  HIP_TRY(c, hipEventCreate(&e1));
  static const uint32_t kRefill[] = {8, 16, 24, 32, 48};
  static const uint32_t kBatch[] = {32, 64, 128};
  int rc = ARC_OK;
  // median of `r` timed launches (after one warm-up) of the shape in *slot
  auto time_shape = [&](uint32_t r, std::vector<float>& t) -> float {
    rc = run();  // warm-up
    for (uint32_t i = 0; i < r && rc == ARC_OK; ++i) {
      if (hipEventRecord(e0, s) != hipSuccess) rc = ARC_ERR_HIP;
      if (rc == ARC_OK) rc = run();
      if (rc == ARC_OK && hipEventRecord(e1, s) != hipSuccess) rc = ARC_ERR_HIP;
      if (rc == ARC_OK && hipEventSynchronize(e1) != hipSuccess) rc = ARC_ERR_HIP;
      if (rc == ARC_OK && hipEventElapsedTime(&t[i], e0, e1) != hipSuccess) rc = ARC_ERR_HIP;
    }
    if (rc != ARC_OK) return 1e30f;
    std::vector<float> sorted(t.begin(), t.begin() + r);
    std::sort(sorted.begin(), sorted.end());
    return sorted[r / 2];
  };
  // Sweep every shape with `reps` launches, then re-time the four fastest with four times as many
  // (at least 32): neighbouring shapes differ by about the launch-to-launch noise, so the short
  // sweep alone can pick a shape 1-2% slower than the best.
  const uint32_t reps2 = std::min(64u, std::max(reps, 8u) * 4u);
  std::vector<std::pair<float, arc_launch_config>> swept;
  std::vector<float> t(reps2);
  for (uint32_t bt : kThreads)
    for (uint32_t bpc : kBlocks)
      for (uint32_t rf : kRefill)
        for (uint32_t bat : kBatch) {
          if (rc != ARC_OK) break;
          if (skip(bt, bpc)) continue;
          *slot = arc_launch_config{bpc, bt, rf, bat};
          const float med = time_shape(reps, t);
          if (rc != ARC_OK) break;
          swept.emplace_back(med, *slot);
        }
  arc_launch_config win = saved;
  float win_ms = 1e30f;
  if (rc == ARC_OK) {
    std::sort(swept.begin(), swept.end(),
              [](const std::pair<float, arc_launch_config>& x, const std::pair<float, arc_launch_config>& y) {
                return x.first < y.first;
              });
    for (size_t k = 0; k < swept.size() && k < 4; ++k) {
      *slot = swept[k].second;
      const float med = time_shape(reps2, t);
      if (rc != ARC_OK) break;
      if (med < win_ms) {
        win_ms = med;
        win = *slot;
      }
    }
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (rc != ARC_OK) {
    *slot = saved;
    return rc;
  }
  *slot = win;
  if (best) *best = win;
  if (best_ms) *best_ms = win_ms;
  return ARC_OK;
}

}  // namespace

extern "C" {

int arc_create(int device, arc_ctx** out) {
  if (!out) return ARC_ERR_INVALID;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) return ARC_ERR_INVALID;
  arc_ctx* c = new (std::nothrow) arc_ctx();
  if (!c) return ARC_ERR_INVALID;
  c->device = device;
  for (auto& k : c->cfg) k = default_config();
  c->cfg_bounce = default_config();
  c->cfg_bounce.blocks_per_cu = 0;
  c->cfg_bounce.refill_lanes = 24;
  c->cfg_bounce.batch_rays = 64;
  c->cfg_shadow = default_config();
  if (hipSetDevice(device) != hipSuccess) {
    delete c;
    return ARC_ERR_HIP;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
    delete c;
    return ARC_ERR_HIP;
  }
  c->n_cu = (uint32_t)prop.multiProcessorCount;
  c->lds_per_block = (uint32_t)prop.sharedMemPerBlock;
  if (hipMalloc(&c->d_queue, kQueueBytes) != hipSuccess) {
    delete c;
    return ARC_ERR_HIP;
  }
  if (hipMemset(c->d_queue, 0, kQueueBytes) != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
    (void)hipFree(c->d_queue);
    delete c;
    return ARC_ERR_HIP;
  }
  float pal[256 * 8];
  for (uint32_t m = 0; m < 256; ++m) arc_palette_material(m, pal + 8 * m);
  if (hipMalloc(&c->d_palette, sizeof pal) != hipSuccess ||
      hipMemcpy(c->d_palette, pal, sizeof pal, hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(c->d_queue);
    (void)hipFree(c->d_palette);
    delete c;
    return ARC_ERR_HIP;
  }
  *out = c;
  return ARC_OK;
}

void arc_destroy(arc_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  (void)hipDeviceSynchronize();
  (void)hipFree(c->d_roots);
  (void)hipFree(c->d_nodes);
  (void)hipFree(c->d_bricks);
  (void)hipFree(c->d_bodies);
  (void)hipFree(c->d_queue);
  (void)hipFree(c->d_msg);
  (void)hipFree(c->d_rays);
  (void)hipFree(c->d_hits);
  (void)hipFree(c->d_palette);
  (void)hipFree(c->d_gb_rays);
  (void)hipFree(c->d_gb_res);
  (void)hipFree(c->d_stage);
  if (c->h_stage) (void)hipHostFree(c->h_stage);
  for (int i = 0; i < 2; ++i) {
    if (c->h_snap[i]) (void)hipHostFree(c->h_snap[i]);
    (void)hipFree(c->d_snap_table[i]);
  }
  (void)hipFree(c->d_restir);
  delete c;
}

const char* arc_last_error(const arc_ctx* c) { return c ? c->err.c_str() : "null context"; }

int arc_device_info(const arc_ctx* c, uint32_t* n_cu, uint32_t* lds) {
  if (!c) return ARC_ERR_INVALID;
  if (n_cu) *n_cu = c->n_cu;
  if (lds) *lds = c->lds_per_block;
  return ARC_OK;
}

static int normalize_config(arc_ctx* c, const arc_launch_config* cfg, arc_launch_config* out,
                            bool path_kernel = false) {
  arc_launch_config d = default_config();
  arc_launch_config n = *cfg;
  if (n.block_threads == 0) n.block_threads = d.block_threads;
  if (n.refill_lanes == 0) n.refill_lanes = d.refill_lanes;
  if (n.batch_rays == 0) n.batch_rays = d.batch_rays;
  const bool large = path_kernel && n.block_threads == arcdev::kPathLargeBlock;
  if (!large && (n.block_threads % 64 != 0 || n.block_threads > 256))
    return fail(c, ARC_ERR_INVALID, path_kernel ? "block_threads must be 64..256 (a multiple of 64) or 1024"
                                                : "block_threads must be 64..256, a multiple of 64");
  if (n.refill_lanes > 64) return fail(c, ARC_ERR_INVALID, "refill_lanes must be 1..64");
  if (n.blocks_per_cu > 64) return fail(c, ARC_ERR_INVALID, "blocks_per_cu must be 0..64");
  if (n.batch_rays > (1u << 20)) return fail(c, ARC_ERR_INVALID, "batch_rays must be 1..2^20");
  *out = n;
  return ARC_OK;
}

int arc_set_launch_config(arc_ctx* c, const arc_launch_config* cfg) {
  if (!c || !cfg) return ARC_ERR_INVALID;
  arc_launch_config n;
  int rc = normalize_config(c, cfg, &n);
  if (rc != ARC_OK) return rc;
  for (auto& k : c->cfg) k = n;
  return ARC_OK;
}

int arc_get_launch_config(const arc_ctx* c, arc_launch_config* cfg) {
  if (!c || !cfg) return ARC_ERR_INVALID;
  *cfg = c->cfg[0];
  return ARC_OK;
}

// Launch-shape slot of a class: the three LOD classes of the traversal, the one-launch frame, the
// shadow walkers.
static arc_launch_config* class_slot(arc_ctx* c, uint32_t cls) {
  if (cls == ARC_LAUNCH_CLASS_BOUNCE) return &c->cfg_bounce;
  if (cls == ARC_LAUNCH_CLASS_SHADOW) return &c->cfg_shadow;
  return cls < ARC_LAUNCH_CLASS_BOUNCE ? &c->cfg[cls] : nullptr;
}

int arc_set_launch_config_class(arc_ctx* c, uint32_t cls, const arc_launch_config* cfg) {
  if (!c || !cfg || !class_slot(c, cls)) return ARC_ERR_INVALID;
  return normalize_config(c, cfg, class_slot(c, cls));
}

int arc_get_launch_config_class(const arc_ctx* c, uint32_t cls, arc_launch_config* cfg) {
  if (!c || !cfg) return ARC_ERR_INVALID;
  const arc_launch_config* slot = class_slot(const_cast<arc_ctx*>(c), cls);
  if (!slot) return ARC_ERR_INVALID;
  *cfg = *slot;
  return ARC_OK;
}

int arc_world_upload(arc_ctx* c, const uint32_t grid[3], const uint32_t* roots, const arc_node64* nodes,
                     uint64_t n_nodes, uint64_t node_cap, const arc_brick64* bricks, uint64_t n_bricks,
                     uint64_t brick_cap) {
  if (!c || !grid || !roots || !nodes || !bricks || n_nodes == 0 || n_bricks == 0)
    return fail(c, ARC_ERR_INVALID, "arc_world_upload: null pointer or empty pool");
  for (int a = 0; a < 3; ++a)
    if (grid[a] == 0 || grid[a] > 256) return fail(c, ARC_ERR_INVALID, "grid dims must be 1..256");
  if (n_nodes > ARC_PTR_INDEX_MASK || n_bricks > ARC_PTR_INDEX_MASK)
    return fail(c, ARC_ERR_INVALID, "pool larger than the 30-bit pointer range");
  if (node_cap < n_nodes) node_cap = n_nodes;
  if (brick_cap < n_bricks) brick_cap = n_bricks;
  const uint64_t nchunks = (uint64_t)grid[0] * grid[1] * grid[2];
  for (uint64_t i = 0; i < nchunks; ++i) {
    if (roots[i] == ARC_EMPTY_CHUNK_SENTINEL) continue;
    std::string why;
    if (arc::check_tree(nodes, n_nodes, n_bricks, roots[i], why) != ARC_OK)
      return fail(c, ARC_ERR_INVALID, "chunk %llu: %s", (unsigned long long)i, why.c_str());
  }
  HIP_TRY(c, hipSetDevice(c->device));
  HIP_TRY(c, hipDeviceSynchronize());
  (void)hipFree(c->d_roots);
  (void)hipFree(c->d_nodes);
  (void)hipFree(c->d_bricks);
  c->d_roots = nullptr;
  c->d_nodes = nullptr;
  c->d_bricks = nullptr;
  c->has_world = false;
  HIP_TRY(c, hipMalloc(&c->d_roots, nchunks * sizeof(uint32_t)));
  HIP_TRY(c, hipMalloc(&c->d_nodes, node_cap * sizeof(arc_node64)));
  HIP_TRY(c, hipMalloc(&c->d_bricks, brick_cap * sizeof(arc_brick64)));
  HIP_TRY(c, hipMemcpy(c->d_roots, roots, nchunks * sizeof(uint32_t), hipMemcpyHostToDevice));
  HIP_TRY(c, hipMemcpy(c->d_nodes, nodes, n_nodes * sizeof(arc_node64), hipMemcpyHostToDevice));
  HIP_TRY(c, hipMemcpy(c->d_bricks, bricks, n_bricks * sizeof(arc_brick64), hipMemcpyHostToDevice));
  if (node_cap > n_nodes) HIP_TRY(c, hipMemset(c->d_nodes + n_nodes, 0, (node_cap - n_nodes) * sizeof(arc_node64)));
  if (brick_cap > n_bricks)
    HIP_TRY(c, hipMemset(c->d_bricks + n_bricks, 0, (brick_cap - n_bricks) * sizeof(arc_brick64)));
  HIP_TRY(c, hipDeviceSynchronize());
  std::memcpy(c->grid, grid, sizeof c->grid);
  c->n_nodes = n_nodes;
  c->node_cap = node_cap;
  c->n_bricks = n_bricks;
  c->brick_cap = brick_cap;
  c->water = arc::pools_hold_material(nodes, n_nodes, bricks, n_bricks, ARC_MAT_WATER);
  c->has_world = true;
  c->world_gen++;
  return ARC_OK;
}

// Streaming pool builder: a snapshot's chunks, in their stored (Morton) order, go straight to
// their final pool offsets. Batches of up to 128 MiB are staged in two pinned buffers: while the
// DMA engine copies one batch, host threads check (CRC, tree rules) and stage the next. A small
// kernel then rebases the batch's node pointers in HBM. The pools equal those of
// arc_world_host_pack followed by arc_world_upload, without building them on the host.
int arc_snapshot_upload(arc_ctx* c, const arc_snapshot* sh, uint64_t node_cap, uint64_t brick_cap, int nthreads) {
  if (!c || !sh) return fail(c, ARC_ERR_INVALID, "arc_snapshot_upload: null argument");
  const arc::Snapshot& s = *reinterpret_cast<const arc::Snapshot*>(sh);
  const arc::SnapHeader& h = s.h;
  const uint64_t n_nodes = h.total_nodes + 1, n_bricks = h.total_bricks + 1;  // slot-0 dummies
  if (node_cap < n_nodes) node_cap = n_nodes;
  if (brick_cap < n_bricks) brick_cap = n_bricks;
  if (node_cap > (uint64_t)ARC_PTR_INDEX_MASK + 1 || brick_cap > (uint64_t)ARC_PTR_INDEX_MASK + 1)
    return fail(c, ARC_ERR_INVALID, "pool capacity beyond the 30-bit pointer range");
  const uint64_t nchunks = (uint64_t)h.grid[0] * h.grid[1] * h.grid[2];
  // batches of consecutive chunks
  struct Batch {
    uint32_t k0, k1;
    uint64_t node0, brick0, n_nodes, n_bricks;
  };
  constexpr uint64_t kBatchBytes = 128ull << 20;
  std::vector<Batch> batches;
  std::vector<uint32_t> roots(nchunks, ARC_EMPTY_CHUNK_SENTINEL);
  std::vector<uint64_t> node_at(h.n_present), brick_at(h.n_present);
  uint64_t stage_max = 0, nb = 1, bb = 1;
  uint32_t table_max = 0;
  for (uint32_t k = 0; k < h.n_present; ++k) {
    const arc::SnapChunk& e = s.index[k];
    const uint64_t bytes = arc::snap_payload_bytes(e);
    if (batches.empty() || (batches.back().n_nodes * 16 + batches.back().n_bricks * 64 + bytes > kBatchBytes &&
                            batches.back().k1 > batches.back().k0))
      batches.push_back(Batch{k, k, nb, bb, 0, 0});
    Batch& b = batches.back();
    b.k1 = k + 1;
    b.n_nodes += e.n_nodes;
    b.n_bricks += e.n_bricks;
    stage_max = std::max(stage_max, b.n_nodes * 16 + b.n_bricks * 64);
    table_max = std::max(table_max, b.k1 - b.k0);
    node_at[k] = nb;
    brick_at[k] = bb;
    roots[e.chunk] = (uint32_t)nb;
    nb += e.n_nodes;
    bb += e.n_bricks;
  }
  HIP_TRY(c, hipSetDevice(c->device));
  HIP_TRY(c, hipDeviceSynchronize());
  (void)hipFree(c->d_roots);
  (void)hipFree(c->d_nodes);
  (void)hipFree(c->d_bricks);
  c->d_roots = nullptr;
  c->d_nodes = nullptr;
  c->d_bricks = nullptr;
  c->has_world = false;
  HIP_TRY(c, hipMalloc(&c->d_roots, nchunks * sizeof(uint32_t)));
  HIP_TRY(c, hipMalloc(&c->d_nodes, node_cap * sizeof(arc_node64)));
  HIP_TRY(c, hipMalloc(&c->d_bricks, brick_cap * sizeof(arc_brick64)));
  // staging: the context's pinned buffers (grown when too small) and two events, destroyed on
  // every exit after the copies that read the buffers have finished
  struct Events {
    hipEvent_t ev[2] = {nullptr, nullptr};
    ~Events() {
      for (auto& e : ev)
        if (e) {
          (void)hipEventSynchronize(e);
          (void)hipEventDestroy(e);
        }
    }
  } st;
  const size_t table_bytes = ((size_t)std::max<uint32_t>(table_max, 1u) * sizeof(uint4) + 255) & ~(size_t)255;
  if (!batches.empty()) {
    if (c->snap_bytes < table_bytes + stage_max || c->snap_table_bytes < table_bytes) {
      for (int i = 0; i < 2; ++i) {
        if (c->h_snap[i]) (void)hipHostFree(c->h_snap[i]);
        (void)hipFree(c->d_snap_table[i]);
        c->h_snap[i] = nullptr;
        c->d_snap_table[i] = nullptr;
      }
      c->snap_bytes = c->snap_table_bytes = 0;
      for (int i = 0; i < 2; ++i) {
        HIP_TRY(c, hipHostMalloc(reinterpret_cast<void**>(&c->h_snap[i]), table_bytes + stage_max, hipHostMallocDefault));
        HIP_TRY(c, hipMalloc(&c->d_snap_table[i], table_bytes));
      }
      c->snap_bytes = table_bytes + stage_max;
      c->snap_table_bytes = table_bytes;
    }
    for (int i = 0; i < 2; ++i) HIP_TRY(c, hipEventCreateWithFlags(&st.ev[i], hipEventDisableTiming));
  }
  const hipStream_t stream = nullptr;
  const int nt = std::max(1, std::min(nthreads, 64));
  std::atomic<bool> water{false};  // any staged chunk holds water
  for (size_t bi = 0; bi < batches.size(); ++bi) {
    const Batch& b = batches[bi];
    const int slot = (int)(bi & 1);
    HIP_TRY(c, hipEventSynchronize(st.ev[slot]));  // the DMA of batch bi - 2 has left this buffer
    uint8_t* const tab = c->h_snap[slot];
    uint8_t* const nodes_h = tab + table_bytes;
    uint8_t* const bricks_h = nodes_h + b.n_nodes * 16;
    std::vector<std::string> why(b.k1 - b.k0);
    std::atomic<uint32_t> next{b.k0};
    std::atomic<int> status{ARC_OK};
    auto work = [&]() {
      for (uint32_t k; (k = next.fetch_add(1)) < b.k1 && status.load() == ARC_OK;) {
        if (arc::snap_check_chunk(s, k, why[k - b.k0]) != ARC_OK) {
          status.store(ARC_ERR_CORRUPT);
          return;
        }
        const arc::SnapChunk& e = s.index[k];
        const uint8_t* p = arc::snap_payload(s, k);
        std::memcpy(nodes_h + (node_at[k] - b.node0) * 16, p, (size_t)e.n_nodes * 16);
        std::memcpy(bricks_h + (brick_at[k] - b.brick0) * 64, p + (size_t)e.n_nodes * 16, (size_t)e.n_bricks * 64);
        if (arc::pools_hold_material(reinterpret_cast<const arc_node64*>(p), e.n_nodes,
                                     reinterpret_cast<const arc_brick64*>(p + (size_t)e.n_nodes * 16), e.n_bricks,
                                     ARC_MAT_WATER))
          water.store(true);
        const uint4 t = make_uint4((uint32_t)node_at[k], e.n_nodes, (uint32_t)brick_at[k], 0u);
        std::memcpy(tab + (size_t)(k - b.k0) * sizeof(uint4), &t, sizeof t);
      }
    };
    std::vector<std::thread> th;
    for (int t = 1; t < std::min<int>(nt, (int)(b.k1 - b.k0)); ++t) th.emplace_back(work);
    work();
    for (auto& t : th) t.join();
    if (status.load() != ARC_OK) {
      (void)hipStreamSynchronize(stream);
      (void)hipFree(c->d_roots);
      (void)hipFree(c->d_nodes);
      (void)hipFree(c->d_bricks);
      c->d_roots = nullptr;
      c->d_nodes = nullptr;
      c->d_bricks = nullptr;
      for (auto& w : why)
        if (!w.empty()) return fail(c, ARC_ERR_CORRUPT, "%s: %s", s.path.c_str(), w.c_str());
      return fail(c, ARC_ERR_CORRUPT, "%s: corrupt chunk", s.path.c_str());
    }
    const uint32_t nt_b = b.k1 - b.k0;
    HIP_TRY(c, hipMemcpyAsync(c->d_snap_table[slot], tab, nt_b * sizeof(uint4), hipMemcpyHostToDevice, stream));
    HIP_TRY(c, hipMemcpyAsync(c->d_nodes + b.node0, nodes_h, b.n_nodes * 16, hipMemcpyHostToDevice, stream));
    if (b.n_bricks)
      HIP_TRY(c, hipMemcpyAsync(c->d_bricks + b.brick0, bricks_h, b.n_bricks * 64, hipMemcpyHostToDevice, stream));
    HIP_TRY(c, arcdev::launch_relocate_chunks(c->d_nodes, (uint32_t)b.node0, (uint32_t)(b.node0 + b.n_nodes),
                                              c->d_snap_table[slot], nt_b, stream));
    HIP_TRY(c, hipEventRecord(st.ev[slot], stream));
  }
  HIP_TRY(c, hipMemsetAsync(c->d_nodes, 0, sizeof(arc_node64), stream));
  HIP_TRY(c, hipMemsetAsync(c->d_bricks, 0, sizeof(arc_brick64), stream));
  if (node_cap > n_nodes)
    HIP_TRY(c, hipMemsetAsync(c->d_nodes + n_nodes, 0, (node_cap - n_nodes) * sizeof(arc_node64), stream));
  if (brick_cap > n_bricks)
    HIP_TRY(c, hipMemsetAsync(c->d_bricks + n_bricks, 0, (brick_cap - n_bricks) * sizeof(arc_brick64), stream));
  HIP_TRY(c, hipMemcpy(c->d_roots, roots.data(), nchunks * sizeof(uint32_t), hipMemcpyHostToDevice));
  HIP_TRY(c, hipDeviceSynchronize());
  std::memcpy(c->grid, h.grid, sizeof c->grid);
  c->n_nodes = n_nodes;
  c->node_cap = node_cap;
  c->n_bricks = n_bricks;
  c->brick_cap = brick_cap;
  c->water = water.load();
  c->has_world = true;
  c->world_gen++;
  return ARC_OK;
}

int arc_world_device_pools(const arc_ctx* c, const uint32_t** roots, const arc_node64** nodes,
                           const arc_brick64** bricks) {
  if (!c || !c->has_world) return ARC_ERR_NO_WORLD;
  if (roots) *roots = c->d_roots;
  if (nodes) *nodes = c->d_nodes;
  if (bricks) *bricks = c->d_bricks;
  return ARC_OK;
}

int arc_world_read_pools(arc_ctx* c, uint32_t* roots, arc_node64* nodes, uint64_t node_first, uint64_t n_nodes,
                         arc_brick64* bricks, uint64_t brick_first, uint64_t n_bricks) {
  if (!c) return ARC_ERR_INVALID;
  if (!c->has_world) return fail(c, ARC_ERR_NO_WORLD, "arc_world_read_pools before arc_world_upload");
  if ((n_nodes && !nodes) || (n_bricks && !bricks)) return fail(c, ARC_ERR_INVALID, "null output");
  if (node_first + n_nodes > c->node_cap || brick_first + n_bricks > c->brick_cap)
    return fail(c, ARC_ERR_INVALID, "read range outside the pools");
  HIP_TRY(c, hipSetDevice(c->device));
  HIP_TRY(c, hipDeviceSynchronize());
  const uint64_t nch = (uint64_t)c->grid[0] * c->grid[1] * c->grid[2];
  if (roots) HIP_TRY(c, hipMemcpy(roots, c->d_roots, nch * sizeof(uint32_t), hipMemcpyDeviceToHost));
  if (n_nodes)
    HIP_TRY(c, hipMemcpy(nodes, c->d_nodes + node_first, n_nodes * sizeof(arc_node64), hipMemcpyDeviceToHost));
  if (n_bricks)
    HIP_TRY(c, hipMemcpy(bricks, c->d_bricks + brick_first, n_bricks * sizeof(arc_brick64), hipMemcpyDeviceToHost));
  return ARC_OK;
}

int arc_world_patch(arc_ctx* c, const arc_gpu_patch* patches, uint64_t n) {
  if (!c) return ARC_ERR_INVALID;
  if (!c->has_world) return fail(c, ARC_ERR_NO_WORLD, "arc_world_patch before arc_world_upload");
  if (n == 0) return ARC_OK;
  if (!patches) return fail(c, ARC_ERR_INVALID, "null patches");
  const uint64_t nchunks = (uint64_t)c->grid[0] * c->grid[1] * c->grid[2];
  arcplan::Plan plan;
  std::string err;
  const int rc = arcplan::plan_patches(patches, n, nchunks, c->node_cap, c->brick_cap, plan, err);
  if (rc != ARC_OK) return fail(c, rc, "%s", err.c_str());
  for (uint64_t i = 0; i < n && !c->water; ++i)
    if (arc::pools_hold_material(patches[i].nodes, patches[i].n_nodes, patches[i].bricks, patches[i].n_bricks,
                                 ARC_MAT_WATER)) {
      c->water = true;
      c->world_gen++;  // captured batches were recorded without the water tests: capture again
    }
  HIP_TRY(c, hipSetDevice(c->device));
  // Work queued on any stream may still read the pools.
  HIP_TRY(c, hipDeviceSynchronize());
  // Grow a pool that a patch writes past (renderer.rs:156-223: copy into a larger buffer).
  if (plan.need_nodes > c->node_cap) {
    const uint64_t cap =
        std::min<uint64_t>(plan.need_nodes + (plan.need_nodes >> 2) + 1024, (uint64_t)ARC_PTR_INDEX_MASK + 1);
    arc_node64* nn = nullptr;
    HIP_TRY(c, hipMalloc(&nn, cap * sizeof(arc_node64)));
    HIP_TRY(c, hipMemcpy(nn, c->d_nodes, c->node_cap * sizeof(arc_node64), hipMemcpyDeviceToDevice));
    HIP_TRY(c, hipMemset(nn + c->node_cap, 0, (cap - c->node_cap) * sizeof(arc_node64)));
    HIP_TRY(c, hipDeviceSynchronize());
    (void)hipFree(c->d_nodes);
    c->d_nodes = nn;
    c->node_cap = cap;
    c->world_gen++;
  }
  if (plan.need_bricks > c->brick_cap) {
    const uint64_t cap =
        std::min<uint64_t>(plan.need_bricks + (plan.need_bricks >> 2) + 1024, (uint64_t)ARC_PTR_INDEX_MASK + 1);
    arc_brick64* nb = nullptr;
    HIP_TRY(c, hipMalloc(&nb, cap * sizeof(arc_brick64)));
    HIP_TRY(c, hipMemcpy(nb, c->d_bricks, c->brick_cap * sizeof(arc_brick64), hipMemcpyDeviceToDevice));
    HIP_TRY(c, hipMemset(nb + c->brick_cap, 0, (cap - c->brick_cap) * sizeof(arc_brick64)));
    HIP_TRY(c, hipDeviceSynchronize());
    (void)hipFree(c->d_bricks);
    c->d_bricks = nb;
    c->brick_cap = cap;
    c->world_gen++;
  }
  const size_t bytes = plan.total_units * 16;
  if (bytes > c->stage_bytes) {
    if (c->h_stage) (void)hipHostFree(c->h_stage);
    if (c->d_stage) (void)hipFree(c->d_stage);
    c->h_stage = nullptr;
    c->d_stage = nullptr;
    c->stage_bytes = 0;
    const size_t cap = bytes + bytes / 4 + 65536;
    HIP_TRY(c, hipHostMalloc(&c->h_stage, cap, hipHostMallocDefault));
    HIP_TRY(c, hipMalloc(&c->d_stage, cap));
    c->stage_bytes = cap;
  }
  arcplan::fill_staging(plan, c->h_stage);
  HIP_TRY(c, hipMemcpy(c->d_stage, c->h_stage, bytes, hipMemcpyHostToDevice));
  const uint4* d = static_cast<const uint4*>(c->d_stage);
  const uint32_t np = (uint32_t)plan.pieces.size();
  hipError_t e = arcdev::launch_apply_patches(reinterpret_cast<const arcdev::PatchPiece*>(d), np,
                                              reinterpret_cast<const uint2*>(d + np), (uint32_t)plan.roots.size(), d,
                                              c->d_nodes, c->d_bricks, c->d_roots, nullptr);
  if (e != hipSuccess) return fail(c, ARC_ERR_HIP, "patch scatter launch: %s", hipGetErrorString(e));
  HIP_TRY(c, hipDeviceSynchronize());
  c->n_nodes = std::max(c->n_nodes, plan.max_node_written);
  c->n_bricks = std::max(c->n_bricks, plan.max_brick_written);
  return ARC_OK;
}

// The device pools (at their capacities, which patches fill up to) each span at most 4 GiB: the
// traversal kernels then address them with 32-bit byte offsets.
static bool pools_off32(const arc_ctx* c) {
  return c->node_cap * sizeof(arc_node64) <= (1ull << 32) && c->brick_cap * sizeof(arc_brick64) <= (1ull << 32);
}

static int launch(arc_ctx* c, const arc_trace_params* params, const arc_ray* d_rays, uint64_t n, void* d_out, bool full,
                  void* stream) {
  if (!c) return ARC_ERR_INVALID;
  if (!valid_params(params)) return fail(c, ARC_ERR_INVALID, "invalid trace params");
  if (!c->has_world) return fail(c, ARC_ERR_NO_WORLD, "trace before arc_world_upload");
  if (n == 0) return ARC_OK;
  if (!d_rays || !d_out) return fail(c, ARC_ERR_INVALID, "null ray or output pointer");
  if (n > 0x7FFFFFFFull) return fail(c, ARC_ERR_INVALID, "at most 2^31-1 rays per launch");
  arcdev::LaunchArgs a{};
  std::memcpy(a.grid, c->grid, sizeof a.grid);
  a.roots = c->d_roots;
  a.nodes = c->d_nodes;
  a.bricks = c->d_bricks;
  a.params = *params;
  if (!c->water) a.params.ignore_water = 0u;  // no water in the pools: the same results without its tests
  a.rays = d_rays;
  a.out = d_out;
  a.n = n;
  a.full = full;
  a.off32 = pools_off32(c);
  a.queue = c->d_queue;
  const arc_launch_config& cfg = c->cfg[lod_class(params)];
  uint32_t bpc = cfg.blocks_per_cu;
  if (bpc == 0) {
    hipError_t oe = arcdev::kernel_occupancy(params->semantics, full, cfg.block_threads, &bpc);
    if (oe != hipSuccess) return fail(c, ARC_ERR_HIP, "occupancy query: %s", hipGetErrorString(oe));
  }
  uint64_t want_blocks = (uint64_t)bpc * c->n_cu;
  uint64_t max_useful = (n + cfg.block_threads - 1) / cfg.block_threads;
  a.blocks = (uint32_t)(want_blocks < max_useful ? want_blocks : max_useful);
  if (a.blocks == 0) a.blocks = 1;
  a.block_threads = cfg.block_threads;
  a.refill = cfg.refill_lanes;
  a.batch = cfg.batch_rays;
  hipError_t e = arcdev::launch_trace(a, (hipStream_t)stream);
  if (e != hipSuccess) return fail(c, ARC_ERR_HIP, "trace launch: %s", hipGetErrorString(e));
  return ARC_OK;
}

int arc_trace_shadow_device(arc_ctx* c, const arc_shadow_params* sp, const arc_ray* d_rays, uint64_t n, float* d_vis,
                            arc_shadow_rec* d_rec, void* stream) {
  if (!c || !sp) return ARC_ERR_INVALID;
  if (sp->mode != ARC_SHADOW_SUN && sp->mode != ARC_SHADOW_DI) return fail(c, ARC_ERR_INVALID, "unknown shadow mode");
  if (!c->has_world) return fail(c, ARC_ERR_NO_WORLD, "trace before arc_world_upload");
  if (n == 0) return ARC_OK;
  if (!d_rays || !d_vis) return fail(c, ARC_ERR_INVALID, "null ray or visibility pointer");
  if (n > 0x7FFFFFFFull) return fail(c, ARC_ERR_INVALID, "at most 2^31-1 rays per launch");
  arcdev::ShadowLaunchArgs a{};
  std::memcpy(a.grid, c->grid, sizeof a.grid);
  a.roots = c->d_roots;
  a.nodes = c->d_nodes;
  a.bricks = c->d_bricks;
  a.mode = sp->mode;
  a.max_steps = sp->max_steps;
  a.rays = d_rays;
  a.vis = d_vis;
  a.rec = d_rec;
  a.n = n;
  a.queue = c->d_queue + 2u * arcdev::kQueueRegionWords;  // the walkers' own queue
  // The walkers have their own shape (ARC_LAUNCH_CLASS_SHADOW, arc_autotune_shadow): the traversal
  // classes' shapes, tuned for long rays, measured slower on these short rays (sun 0.220 vs 0.199 ms).
  const arc_launch_config cfg = c->cfg_shadow;
  uint32_t bpc = cfg.blocks_per_cu;
  if (bpc == 0) {
    hipError_t oe = arcdev::shadow_occupancy(cfg.block_threads, &bpc);
    if (oe != hipSuccess) return fail(c, ARC_ERR_HIP, "occupancy query: %s", hipGetErrorString(oe));
  }
  const uint64_t want_blocks = (uint64_t)bpc * c->n_cu;
  const uint64_t max_useful = (n + cfg.block_threads - 1) / cfg.block_threads;
  a.blocks = (uint32_t)std::max<uint64_t>(1u, std::min(want_blocks, max_useful));
  a.block_threads = cfg.block_threads;
  a.refill = cfg.refill_lanes;
  a.batch = cfg.batch_rays;
  hipError_t e = arcdev::launch_shadow(a, (hipStream_t)stream);
  if (e != hipSuccess) return fail(c, ARC_ERR_HIP, "shadow launch: %s", hipGetErrorString(e));
  return ARC_OK;
}

int arc_autotune_shadow(arc_ctx* c, const arc_shadow_params* sp, const arc_ray* d_rays, uint64_t n, float* d_vis,
                        arc_shadow_rec* d_rec, void* stream, uint32_t reps, arc_launch_config* best, float* best_ms) {
  if (!c) return ARC_ERR_INVALID;
  if (!sp || n == 0 || !d_rays || !d_vis) return fail(c, ARC_ERR_INVALID, "autotune needs a non-empty ray set");
  return tune_shapes(
      c, &c->cfg_shadow, stream, reps, [&] { return arc_trace_shadow_device(c, sp, d_rays, n, d_vis, d_rec, stream); },
      best, best_ms, {128u, 256u}, {0u, 4u, 6u}, [](uint32_t, uint32_t) { return false; });
}

int arc_trace_rigid_device(arc_ctx* c, const arc_rigid_body* bodies, uint32_t n_bodies, const arc_ray* d_rays,
                           uint64_t n, arc_rigid_hit* d_out, void* stream) {
  if (!c) return ARC_ERR_INVALID;
  if (!c->has_world) return fail(c, ARC_ERR_NO_WORLD, "trace before arc_world_upload");
  if (n == 0) return ARC_OK;
  if (!d_rays || !d_out || (n_bodies && !bodies)) return fail(c, ARC_ERR_INVALID, "null ray, output or body pointer");
  if (n > 0x7FFFFFFFull) return fail(c, ARC_ERR_INVALID, "at most 2^31-1 rays per launch");
  HIP_TRY(c, hipSetDevice(c->device));
  const uint32_t nb = std::min<uint32_t>(n_bodies, ARC_MAX_RIGID_BODIES);  // trace_rigid_bodies :1219
  hipStream_t s = (hipStream_t)stream;
  if (nb) {
    if (!c->d_bodies) HIP_TRY(c, hipMalloc(&c->d_bodies, ARC_MAX_RIGID_BODIES * sizeof(arc_rigid_body)));
    HIP_TRY(c, hipMemcpyAsync(c->d_bodies, bodies, nb * sizeof(arc_rigid_body), hipMemcpyHostToDevice, s));
  }
  arcdev::RigidLaunchArgs a{};
  a.nodes = c->d_nodes;
  a.bricks = c->d_bricks;
  a.n_nodes = c->n_nodes;
  a.n_bricks = c->n_bricks;
  a.bodies = c->d_bodies;
  a.n_bodies = nb;
  a.rays = d_rays;
  a.n = n;
  a.out = d_out;
  hipError_t e = arcdev::launch_rigid(a, s);
  if (e != hipSuccess) return fail(c, ARC_ERR_HIP, "rigid launch: %s", hipGetErrorString(e));
  return ARC_OK;
}

int arc_query_voxels_device(arc_ctx* c, const float* d_points_m, uint64_t n, uint32_t* d_mat, void* stream) {
  if (!c) return ARC_ERR_INVALID;
  if (!c->has_world) return fail(c, ARC_ERR_NO_WORLD, "query before arc_world_upload");
  if (n == 0) return ARC_OK;
  if (!d_points_m || !d_mat) return fail(c, ARC_ERR_INVALID, "null point or output pointer");
  if (n > 0x7FFFFFFFull) return fail(c, ARC_ERR_INVALID, "at most 2^31-1 points per launch");
  arcdev::QueryLaunchArgs a{};
  std::memcpy(a.grid, c->grid, sizeof a.grid);
  a.roots = c->d_roots;
  a.nodes = c->d_nodes;
  a.bricks = c->d_bricks;
  a.n_nodes = c->n_nodes;
  a.n_bricks = c->n_bricks;
  a.points = d_points_m;
  a.n = n;
  a.mat = d_mat;
  hipError_t e = arcdev::launch_query(a, (hipStream_t)stream);
  if (e != hipSuccess) return fail(c, ARC_ERR_HIP, "query launch: %s", hipGetErrorString(e));
  return ARC_OK;
}

int arc_trace_device(arc_ctx* c, const arc_trace_params* params, const arc_ray* d_rays, uint64_t n, arc_hit* d_hits,
                     void* stream) {
  return launch(c, params, d_rays, n, d_hits, false, stream);
}

int arc_trace_full_device(arc_ctx* c, const arc_trace_params* params, const arc_ray* d_rays, uint64_t n,
                          arc_trace_result* d_out, void* stream) {
  return launch(c, params, d_rays, n, d_out, true, stream);
}

int arc_trace_pools_noalloc(const uint32_t grid[3], const uint32_t* d_roots, const arc_node64* d_nodes,
                            const arc_brick64* d_bricks, const arc_trace_params* params, const arc_ray* d_rays,
                            uint64_t n, arc_hit* d_hits, void* stream) {
  if (!grid || !valid_params(params)) return ARC_ERR_INVALID;
  for (int a = 0; a < 3; ++a)
    if (grid[a] == 0 || grid[a] > 256) return ARC_ERR_INVALID;
  if (n == 0) return ARC_OK;
  if (!d_roots || !d_nodes || !d_bricks || !d_rays || !d_hits) return ARC_ERR_INVALID;
  if (n > 0x7FFFFFFFull) return ARC_ERR_CAPACITY;
  uint32_t* queue = nullptr;
  uint32_t n_cu = 0;
  if (arcdev::noalloc_scratch(&queue, &n_cu) != hipSuccess) return ARC_ERR_HIP;
  const arc_launch_config cfg = default_config();
  arcdev::LaunchArgs a{};
  std::memcpy(a.grid, grid, sizeof a.grid);
  a.roots = d_roots;
  a.nodes = d_nodes;
  a.bricks = d_bricks;
  a.params = *params;
  a.rays = d_rays;
  a.out = d_hits;
  a.n = n;
  a.full = false;
  a.queue = queue;
  const uint64_t want_blocks = (uint64_t)cfg.blocks_per_cu * n_cu;
  const uint64_t max_useful = (n + cfg.block_threads - 1) / cfg.block_threads;
  a.blocks = (uint32_t)std::max<uint64_t>(1u, std::min(want_blocks, max_useful));
  a.block_threads = cfg.block_threads;
  a.refill = cfg.refill_lanes;
  a.batch = cfg.batch_rays;
  return arcdev::launch_trace(a, (hipStream_t)stream) == hipSuccess ? ARC_OK : ARC_ERR_HIP;
}

int arc_autotune(arc_ctx* c, const arc_trace_params* params, const arc_ray* d_rays, uint64_t n, arc_hit* d_hits,
                 void* stream, uint32_t reps, arc_launch_config* best, float* best_ms) {
  if (!c) return ARC_ERR_INVALID;
  if (!valid_params(params)) return fail(c, ARC_ERR_INVALID, "invalid trace params");
  if (n == 0 || !d_rays || !d_hits) return fail(c, ARC_ERR_INVALID, "autotune needs a non-empty ray set");
  return tune_shapes(
      c, &c->cfg[lod_class(params)], stream, reps, [&] { return launch(c, params, d_rays, n, d_hits, false, stream); },
      best, best_ms, {128u, 256u}, {0u, 4u, 6u}, [](uint32_t, uint32_t) { return false; });
}

// ---- primary rays and their bounces in one launch ------------------------------------------
int arc_trace_bounce_device(arc_ctx* c, const arc_trace_params* primary, const arc_trace_params* bounce,
                            const arc_ray* d_rays, uint64_t n, uint64_t seed, float bounce_max_dist_m,
                            arc_hit* d_hits, void* stream) {
  if (!c) return ARC_ERR_INVALID;
  if (!valid_params(primary) || !valid_params(bounce) || primary->semantics != ARC_SEM_WORLD ||
      bounce->semantics != ARC_SEM_WORLD)
    return fail(c, ARC_ERR_INVALID, "arc_trace_bounce_device traces world semantics (both classes)");
  if (!c->has_world) return fail(c, ARC_ERR_NO_WORLD, "trace before arc_world_upload");
  if (n == 0) return ARC_OK;
  if (!d_rays || !d_hits) return fail(c, ARC_ERR_INVALID, "null ray or output pointer");
  if (n > 0x7FFFFFFFull) return fail(c, ARC_ERR_INVALID, "at most 2^31-1 rays per launch");
  HIP_TRY(c, hipSetDevice(c->device));
  if (n > c->msg_n) {
    // the message words start at zero and every launch leaves them zero; grown (and zeroed) here
    // only, outside any launch of the queue's stream order
    HIP_TRY(c, hipDeviceSynchronize());
    (void)hipFree(c->d_msg);
    c->d_msg = nullptr;
    c->msg_n = 0;
    HIP_TRY(c, hipMalloc(&c->d_msg, n * sizeof(uint64_t)));
    HIP_TRY(c, hipMemset(c->d_msg, 0, n * sizeof(uint64_t)));
    HIP_TRY(c, hipDeviceSynchronize());
    c->msg_n = n;
  }
  arcdev::BounceLaunchArgs a{};
  std::memcpy(a.grid, c->grid, sizeof a.grid);
  a.roots = c->d_roots;
  a.nodes = c->d_nodes;
  a.bricks = c->d_bricks;
  a.primary = *primary;
  a.bounce = *bounce;
  if (!c->water) a.primary.ignore_water = a.bounce.ignore_water = 0u;  // no water: same results without its tests
  a.rays = d_rays;
  a.n = n;
  a.seed = seed;
  a.bounce_max_dist_m = bounce_max_dist_m;
  a.out = d_hits;
  a.msg = c->d_msg;
  a.off32 = pools_off32(c);
  a.queue = c->d_queue + 3u * arcdev::kQueueRegionWords;  // two regions of its own
  const arc_launch_config& cfg = c->cfg_bounce;
  uint32_t bpc = cfg.blocks_per_cu;
  if (bpc == 0) {
    hipError_t oe = arcdev::bounce_occupancy(cfg.block_threads, &bpc);
    if (oe != hipSuccess) return fail(c, ARC_ERR_HIP, "occupancy query: %s", hipGetErrorString(oe));
  }
  const uint64_t want_blocks = (uint64_t)bpc * c->n_cu;
  const uint64_t max_useful = (n + cfg.block_threads - 1) / cfg.block_threads;
  a.blocks = (uint32_t)std::max<uint64_t>(1u, std::min(want_blocks, max_useful));
  a.block_threads = cfg.block_threads;
  a.refill = cfg.refill_lanes;
  a.batch = cfg.batch_rays;
  hipError_t e = arcdev::launch_trace_bounce(a, (hipStream_t)stream);
  if (e != hipSuccess) return fail(c, ARC_ERR_HIP, "trace+bounce launch: %s", hipGetErrorString(e));
  return ARC_OK;
}

int arc_autotune_bounce(arc_ctx* c, const arc_trace_params* primary, const arc_trace_params* bounce,
                        const arc_ray* d_rays, uint64_t n, uint64_t seed, float bounce_max_dist_m, arc_hit* d_hits,
                        void* stream, uint32_t reps, arc_launch_config* best, float* best_ms) {
  if (!c) return ARC_ERR_INVALID;
  if (n == 0 || !d_rays || !d_hits) return fail(c, ARC_ERR_INVALID, "autotune needs a non-empty ray set");
  return tune_shapes(
      c, &c->cfg_bounce, stream, reps,
      [&] { return arc_trace_bounce_device(c, primary, bounce, d_rays, n, seed, bounce_max_dist_m, d_hits, stream); },
      best, best_ms, {128u, 256u}, {0u, 4u, 6u}, [](uint32_t, uint32_t) { return false; });
}

int arc_primary_gbuffer(arc_ctx* c, const arc_trace_params* params, const arc_gbuffer_params* g, uint32_t* d_depth_mat,
                        uint32_t* d_norm_motion, uint32_t* d_albedo_ao, void* stream) {
  if (!c) return ARC_ERR_INVALID;
  if (!g || !params || !d_depth_mat || !d_norm_motion || !d_albedo_ao)
    return fail(c, ARC_ERR_INVALID, "arc_primary_gbuffer: null argument");
  if (g->width == 0 || g->height == 0 || g->width > 16384 || g->height > 16384)
    return fail(c, ARC_ERR_INVALID, "image size must be 1..16384 per axis");
  if (params->semantics != ARC_SEM_WORLD) return fail(c, ARC_ERR_INVALID, "the G-buffer pass traces world semantics");
  if (!c->has_world) return fail(c, ARC_ERR_NO_WORLD, "trace before arc_world_upload");
  HIP_TRY(c, hipSetDevice(c->device));
  const uint64_t tiles = (uint64_t)((g->width + 7) / 8) * ((g->height + 7) / 8);
  const uint64_t n = tiles * 64u;
  if (n > c->gb_n) {
    HIP_TRY(c, hipStreamSynchronize((hipStream_t)stream));
    (void)hipFree(c->d_gb_rays);
    (void)hipFree(c->d_gb_res);
    c->d_gb_rays = nullptr;
    c->d_gb_res = nullptr;
    c->gb_n = 0;
    HIP_TRY(c, hipMalloc(&c->d_gb_rays, n * sizeof(arc_ray)));
    HIP_TRY(c, hipMalloc(&c->d_gb_res, n * sizeof(arc_trace_result)));
    c->gb_n = n;
  }
  arcdev::GbufArgs a;
  std::memcpy(a.eye, g->eye, sizeof a.eye);
  std::memcpy(a.forward, g->forward, sizeof a.forward);
  std::memcpy(a.right, g->right, sizeof a.right);
  std::memcpy(a.up, g->up, sizeof a.up);
  a.tan_half_fov = g->tan_half_fov;
  a.aspect = g->aspect;
  a.jitter[0] = g->jitter[0];
  a.jitter[1] = g->jitter[1];
  std::memcpy(a.prev_view_proj, g->prev_view_proj, sizeof a.prev_view_proj);
  a.width = g->width;
  a.height = g->height;
  a.max_dist_m = g->max_dist_m;
  a.palette = c->d_palette;
  hipError_t e = arcdev::launch_gbuffer_raygen(a, c->d_gb_rays, (uint32_t)n, (hipStream_t)stream);
  if (e != hipSuccess) return fail(c, ARC_ERR_HIP, "raygen launch: %s", hipGetErrorString(e));
  int rc = launch(c, params, c->d_gb_rays, n, c->d_gb_res, true, stream);
  if (rc != ARC_OK) return rc;
  e = arcdev::launch_gbuffer_pack(a, c->d_gb_rays, c->d_gb_res, (uint32_t)n, d_depth_mat, d_norm_motion, d_albedo_ao,
                                  (hipStream_t)stream);
  if (e != hipSuccess) return fail(c, ARC_ERR_HIP, "G-buffer pack launch: %s", hipGetErrorString(e));
  return ARC_OK;
}

int arc_trace(arc_ctx* c, const arc_trace_params* params, const arc_ray* rays, uint64_t n, arc_hit* hits_out) {
  if (!c) return ARC_ERR_INVALID;
  if (n == 0) return ARC_OK;
  if (!rays || !hits_out) return fail(c, ARC_ERR_INVALID, "null ray or hit pointer");
  HIP_TRY(c, hipSetDevice(c->device));
  if (n > c->scratch_n) {
    (void)hipFree(c->d_rays);
    (void)hipFree(c->d_hits);
    c->d_rays = nullptr;
    c->d_hits = nullptr;
    c->scratch_n = 0;
    HIP_TRY(c, hipMalloc(&c->d_rays, n * sizeof(arc_ray)));
    HIP_TRY(c, hipMalloc(&c->d_hits, n * sizeof(arc_hit)));
    c->scratch_n = n;
  }
  HIP_TRY(c, hipMemcpy(c->d_rays, rays, n * sizeof(arc_ray), hipMemcpyHostToDevice));
  int rc = launch(c, params, c->d_rays, n, c->d_hits, false, nullptr);
  if (rc != ARC_OK) return rc;
  HIP_TRY(c, hipDeviceSynchronize());
  HIP_TRY(c, hipMemcpy(hits_out, c->d_hits, n * sizeof(arc_hit), hipMemcpyDeviceToHost));
  return ARC_OK;
}

// ---- batched views: one persistent path launch ------------------------------------------------
}  // extern "C"

struct arc_batch {
  arc_ctx* ctx = nullptr;
  arcdev::ViewDev* d_views = nullptr;
  uint32_t* d_first = nullptr;
  std::vector<uint64_t> first;  // host copy, n_views + 1
  uint32_t n_views = 0;
  uint64_t n_items = 0;
  uint32_t n_bounces = 0;
  float bounce_tmax = 0.0f;
  uint64_t seed = 0;
  arc_trace_params primary{}, bounce{};
  arc_launch_config cfg{};
  // captured graph of one launch
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t cap_stream = nullptr;
  uint64_t graph_gen = 0;
};

namespace {

int batch_launch(arc_batch* b, arc_hit* d_hits, hipStream_t s) {
  arc_ctx* c = b->ctx;
  if (!c->has_world) return fail(c, ARC_ERR_NO_WORLD, "trace before arc_world_upload");
  arcdev::PathLaunchArgs a{};
  std::memcpy(a.grid, c->grid, sizeof a.grid);
  a.roots = c->d_roots;
  a.nodes = c->d_nodes;
  a.bricks = c->d_bricks;
  a.primary = b->primary;
  a.bounce = b->bounce;
  if (!c->water) a.primary.ignore_water = a.bounce.ignore_water = 0u;  // as in launch()
  a.views = b->d_views;
  a.view_first = b->d_first;
  a.n_views = b->n_views;
  a.n_items = (uint32_t)b->n_items;
  a.n_bounces = b->n_bounces;
  a.bounce_max_dist_m = b->bounce_tmax;
  a.seed = b->seed;
  a.out = d_hits;
  a.off32 = pools_off32(c);
  a.queue = c->d_queue + arcdev::kQueueRegionWords;  // the path kernel's own queue
  uint32_t bpc = b->cfg.blocks_per_cu;
  if (bpc == 0) {
    hipError_t oe = arcdev::path_occupancy(b->cfg.block_threads, &bpc);
    if (oe != hipSuccess) return fail(c, ARC_ERR_HIP, "occupancy query: %s", hipGetErrorString(oe));
  }
  a.block_threads = b->cfg.block_threads;
  a.blocks = std::max<uint32_t>(1u, c->n_cu * bpc);
  a.refill = std::max<uint32_t>(1u, std::min<uint32_t>(b->cfg.refill_lanes, 64u));
  a.batch = std::max<uint32_t>(1u, b->cfg.batch_rays);
  hipError_t e = arcdev::launch_path(a, s);
  if (e != hipSuccess) return fail(c, ARC_ERR_HIP, "path kernel launch: %s", hipGetErrorString(e));
  return ARC_OK;
}

}  // namespace

extern "C" {

int arc_batch_create(arc_ctx* c, const arc_view* views, uint32_t n_views, const arc_trace_params* primary,
                     const arc_trace_params* bounce, uint32_t n_bounces, float bounce_max_dist_m, uint64_t seed,
                     arc_batch** out) {
  if (!c || !out) return ARC_ERR_INVALID;
  *out = nullptr;
  if (!views || n_views == 0 || n_views > 65536) return fail(c, ARC_ERR_INVALID, "1..65536 views");
  if (!valid_params(primary) || !valid_params(bounce) || primary->semantics != ARC_SEM_WORLD ||
      bounce->semantics != ARC_SEM_WORLD)
    return fail(c, ARC_ERR_INVALID, "batches trace world semantics");
  if (n_bounces > 4) return fail(c, ARC_ERR_INVALID, "at most 4 bounces");
  std::vector<arcdev::ViewDev> hv(n_views);
  std::vector<uint32_t> first(n_views + 1);
  uint64_t items = 0;
  for (uint32_t v = 0; v < n_views; ++v) {
    const arc_view& iv = views[v];
    if (iv.width == 0 || iv.height == 0 || iv.width > 16384 || iv.height > 16384)
      return fail(c, ARC_ERR_INVALID, "view %u: size must be 1..16384 per axis", v);
    hv[v].cam = iv.cam;
    hv[v].width = iv.width;
    hv[v].height = iv.height;
    hv[v].tiles_x = (iv.width + 7) / 8;
    hv[v].max_dist_m = iv.max_dist_m;
    first[v] = (uint32_t)items;
    items += (uint64_t)hv[v].tiles_x * ((iv.height + 7) / 8) * 64u;
    if (items * (1u + n_bounces) > 0x7FFFFFFFull) return fail(c, ARC_ERR_CAPACITY, "batch too large");
  }
  first[n_views] = (uint32_t)items;
  arc_batch* b = new (std::nothrow) arc_batch();
  if (!b) return ARC_ERR_CAPACITY;
  b->ctx = c;
  b->n_views = n_views;
  b->n_items = items;
  b->n_bounces = n_bounces;
  b->bounce_tmax = bounce_max_dist_m;
  b->seed = seed;
  b->primary = *primary;
  b->bounce = *bounce;
  b->cfg = default_config();
  b->cfg.blocks_per_cu = 0;
  b->cfg.refill_lanes = 24;
  // Workgroups of 256: on the 4096^3 batched world (16^3 chunks) the 1024-lane shape, which holds
  // every chunk root in LDS, measured 18.8 ms per frame against 17.5 ms (5 waves per SIMD against
  // 4); the autotuner still tries it.
  b->first.assign(first.begin(), first.end());
  hipError_t e = hipSetDevice(c->device);
  if (e == hipSuccess) e = hipMalloc(&b->d_views, n_views * sizeof(arcdev::ViewDev));
  if (e == hipSuccess) e = hipMalloc(&b->d_first, (n_views + 1) * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMemcpy(b->d_views, hv.data(), n_views * sizeof(arcdev::ViewDev), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(b->d_first, first.data(), (n_views + 1) * sizeof(uint32_t), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(b->d_views);
    (void)hipFree(b->d_first);
    delete b;
    return fail(c, ARC_ERR_HIP, "batch upload: %s", hipGetErrorString(e));
  }
  *out = b;
  return ARC_OK;
}

uint64_t arc_batch_slots(const arc_batch* b) { return b ? b->n_items : 0; }

int arc_batch_set_seed(arc_batch* b, uint64_t seed) {
  if (!b) return ARC_ERR_INVALID;
  b->seed = seed;
  return ARC_OK;
}

uint64_t arc_batch_view_first(const arc_batch* b, uint32_t v) {
  return (b && v <= b->n_views) ? b->first[v] : 0;
}

int arc_batch_set_launch_config(arc_batch* b, const arc_launch_config* cfg) {
  if (!b || !cfg) return ARC_ERR_INVALID;
  arc_launch_config n;
  int rc = normalize_config(b->ctx, cfg, &n, true);
  if (rc != ARC_OK) return rc;
  b->cfg = n;
  return ARC_OK;
}

int arc_batch_get_launch_config(const arc_batch* b, arc_launch_config* cfg) {
  if (!b || !cfg) return ARC_ERR_INVALID;
  *cfg = b->cfg;
  return ARC_OK;
}

int arc_batch_trace(arc_batch* b, arc_hit* d_hits, void* stream) {
  if (!b) return ARC_ERR_INVALID;
  if (!d_hits) return fail(b->ctx, ARC_ERR_INVALID, "null hit buffer");
  HIP_TRY(b->ctx, hipSetDevice(b->ctx->device));
  return batch_launch(b, d_hits, (hipStream_t)stream);
}

int arc_batch_capture(arc_batch* b, arc_hit* d_hits) {
  if (!b) return ARC_ERR_INVALID;
  arc_ctx* c = b->ctx;
  if (!d_hits) return fail(c, ARC_ERR_INVALID, "null hit buffer");
  HIP_TRY(c, hipSetDevice(c->device));
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  if (!b->cap_stream) HIP_TRY(c, hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking));
  HIP_TRY(c, hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal));
  const int rc = batch_launch(b, d_hits, b->cap_stream);
  hipGraph_t g = nullptr;
  const hipError_t ee = hipStreamEndCapture(b->cap_stream, &g);
  if (rc != ARC_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (ee != hipSuccess) return fail(c, ARC_ERR_HIP, "graph capture: %s", hipGetErrorString(ee));
  b->graph = g;
  HIP_TRY(c, hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
  b->graph_gen = c->world_gen;
  return ARC_OK;
}

int arc_batch_replay(arc_batch* b, void* stream) {
  if (!b) return ARC_ERR_INVALID;
  arc_ctx* c = b->ctx;
  if (!b->exec) return fail(c, ARC_ERR_INVALID, "arc_batch_replay before arc_batch_capture");
  if (b->graph_gen != c->world_gen)
    return fail(c, ARC_ERR_INVALID, "world pools moved (upload or growth) since the capture: capture again");
  HIP_TRY(c, hipGraphLaunch(b->exec, (hipStream_t)stream));
  return ARC_OK;
}

int arc_batch_autotune(arc_batch* b, arc_hit* d_hits, void* stream, uint32_t reps, arc_launch_config* best,
                       float* best_ms) {
  if (!b) return ARC_ERR_INVALID;
  arc_ctx* c = b->ctx;
  if (!d_hits) return fail(c, ARC_ERR_INVALID, "null hit buffer");
  // 1024 = one workgroup per CU with the chunk roots of up to 4096 chunks in LDS; its LDS fits one
  // workgroup per CU, so it is tried at the occupancy limit only
  return tune_shapes(
      c, &b->cfg, stream, reps, [&] { return batch_launch(b, d_hits, (hipStream_t)stream); }, best, best_ms,
      {128u, 256u, arcdev::kPathLargeBlock}, {0u, 3u, 4u},
      [](uint32_t bt, uint32_t bpc) { return bt == arcdev::kPathLargeBlock && bpc != 0u; });
}

int arc_batch_shade(arc_batch* b, const arc_hit* d_hits, const arc_shade_params* sp, arc_path_sample* d_samples,
                    float* d_radiance, double* d_accum, void* stream) {
  if (!b || !sp) return ARC_ERR_INVALID;
  arc_ctx* c = b->ctx;
  if (!d_hits) return fail(c, ARC_ERR_INVALID, "null hit buffer");
  HIP_TRY(c, hipSetDevice(c->device));
  arcdev::ShadeArgs a{};
  a.views = b->d_views;
  a.view_first = b->d_first;
  a.n_views = b->n_views;
  a.n_items = (uint32_t)b->n_items;
  a.n_bounces = b->n_bounces;
  a.bounce_tmax = b->bounce_tmax;
  a.seed = b->seed;
  a.hits = reinterpret_cast<const uint4*>(d_hits);
  a.palette = c->d_palette;
  std::memcpy(a.sun, sp->sun_dir, sizeof a.sun);
  a.emission_scale = sp->emission_scale;
  a.sky_scale = sp->sky_scale;
  a.samples = d_samples;
  a.radiance = d_radiance;
  a.accum = d_accum;
  hipError_t e = arcdev::launch_shade(a, (hipStream_t)stream);
  if (e != hipSuccess) return fail(c, ARC_ERR_HIP, "shade launch: %s", hipGetErrorString(e));
  return ARC_OK;
}

int arc_estimate(const arc_path_sample* d_samples, const float* d_cache, uint64_t n, uint32_t mode, float* d_out,
                 void* stream) {
  if (n == 0) return ARC_OK;
  if (!d_samples || !d_out || (mode == 1 && !d_cache) || mode > 1 || n > 0xFFFFFFFFull) return ARC_ERR_INVALID;
  return arcdev::launch_estimate(d_samples, d_cache, (uint32_t)n, mode, d_out, (hipStream_t)stream) == hipSuccess
             ? ARC_OK
             : ARC_ERR_HIP;
}

void arc_batch_destroy(arc_batch* b) {
  if (!b) return;
  (void)hipSetDevice(b->ctx->device);
  (void)hipDeviceSynchronize();
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->cap_stream) (void)hipStreamDestroy(b->cap_stream);
  (void)hipFree(b->d_views);
  (void)hipFree(b->d_first);
  delete b;
}


// ---- ReSTIR GI / DI passes ---------------------------------------------------------------------
}  // extern "C"

namespace {

// Carves the context's ReSTIR scratch into the per-pixel buffers of `d` (256-byte aligned).
int restir_scratch(arc_ctx* c, uint64_t n, arcdev::RestirDev& d) {
  const size_t sizes[] = {sizeof(uint2),   sizeof(arc_ray),          sizeof(arc_ray),      sizeof(arc_ray),
                          sizeof(arc_ray), sizeof(arc_hit),          sizeof(float),        sizeof(float),
                          sizeof(arc_trace_result), sizeof(arc_rigid_hit), sizeof(arc_sharc_probe), sizeof(float4),
                          sizeof(float)};
  constexpr int kN = (int)(sizeof sizes / sizeof sizes[0]);
  size_t off[kN], total = 0;
  for (int k = 0; k < kN; ++k) {
    off[k] = total;
    total += (sizes[k] * n + 255) & ~(size_t)255;
  }
  if (total > c->restir_bytes) {
    HIP_TRY(c, hipDeviceSynchronize());  // work queued on any stream may still use the old buffers
    (void)hipFree(c->d_restir);
    c->d_restir = nullptr;
    c->restir_bytes = 0;
    HIP_TRY(c, hipMalloc(&c->d_restir, total));
    c->restir_bytes = total;
  }
  uint8_t* b = c->d_restir;
  d.state = reinterpret_cast<uint2*>(b + off[0]);
  d.rays_a = reinterpret_cast<arc_ray*>(b + off[1]);
  d.rays_b = reinterpret_cast<arc_ray*>(b + off[2]);
  d.rays_c = reinterpret_cast<arc_ray*>(b + off[3]);
  d.rays_d = reinterpret_cast<arc_ray*>(b + off[4]);
  d.hits_a = reinterpret_cast<const arc_hit*>(b + off[5]);
  d.vis_a = reinterpret_cast<const float*>(b + off[6]);
  d.vis_b = reinterpret_cast<const float*>(b + off[7]);
  d.full_b = reinterpret_cast<const arc_trace_result*>(b + off[8]);
  d.rb_c = reinterpret_cast<const arc_rigid_hit*>(b + off[9]);
  d.probes = reinterpret_cast<arc_sharc_probe*>(b + off[10]);
  d.sharc_out = reinterpret_cast<const float4*>(b + off[11]);
  d.vis_d = reinterpret_cast<const float*>(b + off[12]);
  return ARC_OK;
}

int restir_common(arc_ctx* c, const arc_restir_params* p, const uint32_t* dm, const uint32_t* nm,
                  const arc_gi_reservoir* gi_in, arc_gi_reservoir* gi_out, const arc_di_reservoir* di_in,
                  arc_di_reservoir* di_out, const arc_emissive_source* sources, arcdev::RestirDev& d) {
  if (!c || !p) return ARC_ERR_INVALID;
  if (!c->has_world) return fail(c, ARC_ERR_NO_WORLD, "ReSTIR pass before arc_world_upload");
  if (p->width == 0 || p->height == 0 || p->width > 65535 || p->height > 65535)
    return fail(c, ARC_ERR_INVALID, "width and height must be 1..65535 (world_grid_dims.w packs them in 16 bits)");
  if ((uint64_t)p->width * p->height > 0x7FFFFFFFull)
    return fail(c, ARC_ERR_INVALID, "at most 2^31-1 pixels per pass");
  if (!dm || !nm || !gi_in || !gi_out || !di_in || !di_out)
    return fail(c, ARC_ERR_INVALID, "null G-buffer or reservoir pointer");
  if (p->n_emissive > 16384u) return fail(c, ARC_ERR_INVALID, "n_emissive above DI_MAX_SOURCES (16384)");
  if (p->n_emissive && !sources) return fail(c, ARC_ERR_INVALID, "n_emissive > 0 needs the emissive sources");
  if (p->shadow_mode > 1u) return fail(c, ARC_ERR_INVALID, "shadow_mode must be 0 or 1");
  HIP_TRY(c, hipSetDevice(c->device));
  const uint64_t n = (uint64_t)p->width * p->height;
  int rc = restir_scratch(c, n, d);
  if (rc != ARC_OK) return rc;
  d.p = *p;
  d.gx = c->grid[0];
  d.gy = c->grid[1];
  d.gz = c->grid[2];
  d.roots = c->d_roots;
  d.nodes = reinterpret_cast<const uint4*>(c->d_nodes);
  d.palette = c->d_palette;
  d.depth_mat = reinterpret_cast<const uint2*>(dm);
  d.norm_motion = reinterpret_cast<const uint2*>(nm);
  d.gi_in = gi_in;
  d.gi_out = gi_out;
  d.di_in = di_in;
  d.di_out = di_out;
  d.sources = sources;
  return ARC_OK;
}

}  // namespace

extern "C" {

int arc_restir_temporal(arc_ctx* c, const arc_restir_params* p, const arc_trace_params* sun_trace,
                        const arc_trace_params* gi_trace, const uint32_t* d_depth_mat, const uint32_t* d_norm_motion,
                        const arc_gi_reservoir* d_gi_in, arc_gi_reservoir* d_gi_out, const arc_di_reservoir* d_di_in,
                        arc_di_reservoir* d_di_out, const arc_emissive_source* d_sources,
                        const arc_alias_entry* d_alias, const arc_sharc* sharc, const arc_rigid_body* bodies,
                        uint32_t n_bodies, void* stream) {
  if (!sun_trace || !gi_trace) return fail(c, ARC_ERR_INVALID, "null trace parameters");
  if (p && p->n_emissive && !d_alias) return fail(c, ARC_ERR_INVALID, "n_emissive > 0 needs the alias table");
  if (n_bodies && !bodies) return fail(c, ARC_ERR_INVALID, "null rigid bodies");
  arcdev::RestirDev d{};
  int rc = restir_common(c, p, d_depth_mat, d_norm_motion, d_gi_in, d_gi_out, d_di_in, d_di_out, d_sources, d);
  if (rc != ARC_OK) return rc;
  d.alias = d_alias;
  d.has_sharc = sharc ? 1u : 0u;
  d.n_bodies = std::min<uint32_t>(n_bodies, ARC_MAX_RIGID_BODIES);
  const hipStream_t s = (hipStream_t)stream;
  const uint64_t n = (uint64_t)p->width * p->height;
  HIP_TRY(c, arcdev::launch_rt_setup(d, s));
  // sun shadow (:149-159): trace_world_64tree(.., 1, true) or trace_shadow_vis
  arc_shadow_params sun{ARC_SHADOW_SUN, p->shadow_max_steps};
  if (p->shadow_mode == 0u) {
    arc_trace_params tp = *sun_trace;
    tp.semantics = ARC_SEM_WORLD;
    tp.lod_bias = 1u;
    tp.ignore_water = 1u;
    rc = arc_trace_device(c, &tp, d.rays_a, n, const_cast<arc_hit*>(d.hits_a), stream);
  } else {
    rc = arc_trace_shadow_device(c, &sun, d.rays_a, n, const_cast<float*>(d.vis_a), nullptr, stream);
  }
  if (rc != ARC_OK) return rc;
  // GI ray (:207): trace_world_64tree(gi_origin, bounce_dir, gi_max_dist, 2, true)
  arc_trace_params gp = *gi_trace;
  gp.semantics = ARC_SEM_WORLD;
  gp.lod_bias = 2u;
  gp.ignore_water = 1u;
  rc = arc_trace_full_device(c, &gp, d.rays_b, n, const_cast<arc_trace_result*>(d.full_b), stream);
  if (rc != ARC_OK) return rc;
  if (d.n_bodies) {  // :208-215
    HIP_TRY(c, arcdev::launch_rt_rigid_rays(d, s));
    rc = arc_trace_rigid_device(c, bodies, d.n_bodies, d.rays_c, n, const_cast<arc_rigid_hit*>(d.rb_c), stream);
    if (rc != ARC_OK) return rc;
  }
  HIP_TRY(c, arcdev::launch_rt_bounce(d, s));
  if (sharc) {
    rc = arc_sharc_query(sharc, d.probes, n, p->cam.eye, reinterpret_cast<float*>(const_cast<float4*>(d.sharc_out)),
                         stream);
    if (rc != ARC_OK) return fail(c, rc, "SHARC query failed");
  }
  rc = arc_trace_shadow_device(c, &sun, d.rays_d, n, const_cast<float*>(d.vis_d), nullptr, stream);
  if (rc != ARC_OK) return rc;
  HIP_TRY(c, arcdev::launch_rt_resolve(d, s));
  return ARC_OK;
}

int arc_restir_spatial(arc_ctx* c, const arc_restir_params* p, const uint32_t* d_depth_mat,
                       const uint32_t* d_norm_motion, const arc_gi_reservoir* d_gi_in, arc_gi_reservoir* d_gi_out,
                       const arc_di_reservoir* d_di_in, arc_di_reservoir* d_di_out,
                       const arc_emissive_source* d_sources, void* stream) {
  arcdev::RestirDev d{};
  int rc = restir_common(c, p, d_depth_mat, d_norm_motion, d_gi_in, d_gi_out, d_di_in, d_di_out, d_sources, d);
  if (rc != ARC_OK) return rc;
  const hipStream_t s = (hipStream_t)stream;
  const uint64_t n = (uint64_t)p->width * p->height;
  HIP_TRY(c, arcdev::launch_rs_rays(d, s));
  const arc_shadow_params sun{ARC_SHADOW_SUN, p->shadow_max_steps}, di{ARC_SHADOW_DI, p->shadow_max_steps};
  rc = arc_trace_shadow_device(c, &sun, d.rays_a, n, const_cast<float*>(d.vis_a), nullptr, stream);
  if (rc != ARC_OK) return rc;
  rc = arc_trace_shadow_device(c, &di, d.rays_b, n, const_cast<float*>(d.vis_b), nullptr, stream);
  if (rc != ARC_OK) return rc;
  HIP_TRY(c, arcdev::launch_rs_resolve(d, s));
  return ARC_OK;
}

}  // extern "C"
