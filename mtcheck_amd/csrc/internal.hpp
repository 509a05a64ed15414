// internal.hpp -- host-side state shared by the library's translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <atomic>

#include <string>
#include <vector>

#include "../../include/mtcheck.h"
#include "device.hpp"

namespace mtc {

struct HipError {
  hipError_t err;
  const char* what;
};

#define MTC_HIP(call)                                 \
  do {                                                \
    hipError_t e_ = (call);                           \
    if (e_ != hipSuccess) throw ::mtc::HipError{e_, #call}; \
  } while (0)

struct Status {
  int code;
  std::string msg;
};

// Device allocation owned by the library, grown on demand.
template <class T>
struct DevBuf {
  T* p = nullptr;
  uint64_t count = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), count(o.count) {
    o.p = nullptr;
    o.count = 0;
  }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      count = o.count;
      o.p = nullptr;
      o.count = 0;
    }
    return *this;
  }
  ~DevBuf() { release(); }
  void ensure(uint64_t want) {
    if (want <= count) return;
    release();
    MTC_HIP(hipMalloc((void**)&p, (want ? want : 1) * sizeof(T)));
    count = want;
  }
  // As ensure, growing by at least half the current size: for small scratch whose
  // need creeps up call after call (each reallocation synchronises the device).
  void reserve(uint64_t want) {
    if (want <= count) return;
    ensure(std::max<uint64_t>(want, count + count / 2));
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    count = 0;
  }
};

// One fused sweep stage: parents at depths [d_out, d_in-1] over tiles of 2^(d_in-d_out)
// positions at depth d_in.
struct Stage {
  uint32_t d_in, d_out;
};

// Device-resident counters and results of one checkpoint (one 256-byte block).
struct Ctl {
  uint32_t err;
  uint32_t F, S, P;               // roots and prior-checkpoint table entries
  unsigned long long claims;      // new table entries created
  unsigned long long data_chunks; // chunks in the data section
  unsigned long long diff_size;   // bytes of the diff
  unsigned long long changed;     // leaves that are not fixed duplicates
  // Device wall clock (hipDeviceAttributeWallClockRate) when the leaf kernel's
  // first block starts, when the leaf scan's first block starts and when sweep B
  // ends: the phase timers, without event records inside the step.
  unsigned long long t_begin, t_leaf, t_forest;
  unsigned long long t_end;  // ... and when the last kernel writes this record out
  uint32_t pad[44];
};
static_assert(sizeof(Ctl) == 256, "the control block is one 256-byte record");

struct Workspace {
  DevBuf<uint32_t> u32a, u32b, u32c;  // scan partials
};

void exclusive_scan_u32(const uint32_t* in, uint32_t* out, uint64_t count, uint32_t* total, Workspace& ws,
                        hipStream_t st);
// The same scan with caller-provided tile-sum buffers (no allocation): sums[d]
// holds at least scan_sums_words(count, d) words.
uint64_t scan_sums_words(uint64_t count, int depth);
void exclusive_scan_u32_fixed(const uint32_t* in, uint32_t* out, uint64_t count, uint32_t* total,
                              uint32_t* const sums[3], hipStream_t st);

// Launch state of one step (a checkpoint of one or many deduplicators): the
// uploaded argument arena, the step's result array and, when steps are replayed
// from a graph, the captured graph and the launch shape it was captured for.
struct StepCtx {
  DevBuf<uint8_t> arena_d, lead_d;
  uint8_t* arena_h = nullptr;  // pinned staging of the arena of the step's second part
  uint64_t arena_h_cap = 0;
  uint8_t* lead_h = nullptr;  // pinned staging of the arena of its first part (fills, leaf kernels)
  uint64_t lead_h_cap = 0;
  Ctl* ctl_h = nullptr;  // pinned: each member's last kernel writes its record here
  uint32_t ctl_h_cap = 0;
  uint32_t* cnt_h = nullptr;  // pinned: chunks to insert per member (steps that pause after the leaf kernel)
  uint32_t cnt_h_cap = 0;
  hipEvent_t ev[2] = {};  // step start, set-up done
  int clock_khz = 0;  // device wall clock rate (the control block's stamps)
  hipGraphExec_t exec = nullptr;
  std::vector<uint64_t> sig;
  StepCtx() = default;
  StepCtx(const StepCtx&) = delete;
  StepCtx& operator=(const StepCtx&) = delete;
  ~StepCtx();
  void reset_graph();
};

// Table slot indices are 32-bit and index `capacity` is the overflow slot, so a
// table holds at most 2^31 slots.
constexpr uint64_t kMaxTableSlots = 1ull << 31;
// Most members of one batched step (mtc_checkpoint_batch, a serving lane's step).
constexpr uint32_t kMaxBatch = 1u << 16;
// Slots of a baseline's table for n chunks at a load factor (a power of two).
uint64_t table_slots(uint64_t n, double load_factor);
// Device bytes a deduplicator of n chunks holds at a load factor.
uint64_t state_bytes(uint64_t n, uint32_t chunk_size, double load_factor);

// Schedule parameters (set by mtc_set_tuning or the autotuner).
struct Tune {
  uint64_t leaf_tiles_per_block = 1;  // 256-chunk tiles each leaf block streams
  double load_factor = 0.125;         // table entries per slot at a baseline (capacity = n / load_factor)
  uint32_t fuse = 8;                  // tree levels per sweep stage (1..8)
  uint32_t sweep_tiles_a = 0;         // first-stage tiles per wavefront of sweep A / sweep B (0: the
  uint32_t sweep_tiles_b = 0;         // library's default)
};

}  // namespace mtc

// Opaque handle behind include/mtcheck.h.
struct mtc_dedup {
  int device = 0;
  uint32_t chunk_size = 0;
  int hash_kind = 0;
  int approach = 0;
  uint32_t current_id = 0;
  uint32_t baseline_id = 0;
  bool has_state = false;
  // Every leaf label is L_FIXED (the last step reset the chunks it listed), so the
  // next leaf kernel labels only the chunks it lists.
  bool labels_clean = false;
  uint64_t data_len = 0;
  uint32_t n = 0;
  // power of two; slot index cap is the overflow slot.  Atomic: a serving lane may
  // grow a client's table while the client reads its statistics.
  std::atomic<uint64_t> table_cap{0};
  uint64_t table_size = 0;  // entries (host mirror, refreshed after every checkpoint)
  int debug = 0;
  int compat = 0;  // reproduce the reference's bytes where they depart (mtc_set_compat)

  mtc::DevBuf<mtc::Digest> leaf_dg;   // persistent leaf digests (the tree's leaves)
  mtc::DevBuf<uint8_t> labels;        // leaf labels of the current checkpoint
  mtc::DevBuf<uint32_t> leaf_slot;    // table slot of each non-fixed leaf
  mtc::DevBuf<uint32_t> list;         // chunks to insert, packed per leaf block
  mtc::DevBuf<uint32_t> dense;        // the same, concatenated
  mtc::DevBuf<mtc::Digest> list_dg;   // tree handles: their digests beside the list entries
  mtc::DevBuf<mtc::Digest> dense_dg;  // and concatenated
  mtc::DevBuf<uint32_t> pend;         // listed chunks left for k_fixmin / k_resolve
  mtc::DevBuf<uint8_t> pcnt;          // their number per group of 64 dense entries
  mtc::DevBuf<uint32_t> blk_cnt, blk_off;
  uint64_t blk_zero = 0;             // leading entries of blk_cnt known to be zero (k_compact clears them)
  const void* blk_zero_p = nullptr;  // ... in this allocation
  mtc::DevBuf<uint8_t> tile_flags;    // per stage-0 tile: has-first, then busy
  mtc::DevBuf<mtc::Slot> table;
  mtc::DevBuf<mtc::Slot> table_tmp;   // rehash target
  mtc::DevBuf<uint8_t> fr_lab;        // stage frontiers (label, digest, slot)
  mtc::DevBuf<mtc::Digest> fr_dg;
  mtc::DevBuf<uint32_t> fr_slot;
  mtc::DevBuf<uint64_t> staging;      // emitted roots, indexed by node-id range
  mtc::DevBuf<uint32_t> seg_cnt_f, seg_cnt_s, seg_off_f, seg_off_s, seg_base;
  mtc::DevBuf<uint32_t> first_roots;  // F, ascending node id
  mtc::DevBuf<uint64_t> shift_tmp;    // (slot << 32 | node), ascending node id
  mtc::DevBuf<uint32_t> sort_key[2];
  mtc::DevBuf<uint64_t> sort_val[2];  // (prev << 32 | node)
  mtc::DevBuf<uint32_t> sort_hist;
  mtc::DevBuf<uint32_t> src_hist;     // per source checkpoint
  mtc::DevBuf<uint32_t> ptable;       // P (chkpt, count) pairs
  mtc::DevBuf<uint32_t> region_off;   // F region offsets (chunks), then total
  mtc::DevBuf<uint32_t> tile_lo;      // per 256 diff chunks: the region holding the first
  mtc::DevBuf<uint8_t> diff;          // device diff arena 0
  mtc::DevBuf<uint8_t> diff_alt;      // arena 1, allocated once asynchronous copies are used
  int diff_cur = 0;                   // arena holding the last checkpoint's diff
  bool async_copies = false;          // checkpoints alternate arenas
  hipStream_t copy_st = nullptr;      // asynchronous diff copies
  hipEvent_t copy_done[2] = {nullptr, nullptr};  // after the last copy out of each arena
  bool copy_pending[2] = {false, false};
  const uint8_t* diff_ptr() const { return diff_cur ? diff_alt.p : diff.p; }
  mtc::DevBuf<mtc::Digest> dbg_tree;  // debug: full tree, as the serial reference leaves it
  mtc::DevBuf<uint8_t> dbg_labels;    // debug: all node labels
  mtc::DevBuf<unsigned long long> ctr;  // striped per-checkpoint counters
  mtc::DevBuf<uint32_t> scan_ws;      // partial sums of the step's scans
  mtc::Ctl* ctl_d = nullptr;
  mtc::Ctl* ctl_h = nullptr;          // pinned
  mtc::Workspace ws;
  std::vector<mtc::Stage> stages;
  std::vector<uint32_t> seg_base_h;   // first segment of each depth
  std::vector<uint32_t> seg_geo_h;    // per depth: tiles per segment | log2 positions per tile << 16
  uint32_t plan_fuse = 0, plan_group = 0;  // the levels per stage and segment group planned for
  bool plan_dirty = false;            // seg_base_h not yet on the device
  mtc::StepCtx step;                  // launch state of single checkpoints
  mtc::Tune tune;
  mtc_header last_header{};
  uint64_t last_diff_size = 0;
  float last_ms[5] = {0, 0, 0, 0, 0};  // setup, dedup, collect, copy, leaf kernel
  uint32_t last_leaf_labels = 0;       // the last leaf kernel labelled every chunk
  uint32_t last_leaf_digests = 0;      // the last leaf kernel stored the listed chunks' digests
  std::string last_error;
};
