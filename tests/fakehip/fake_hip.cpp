// fake_hip.cpp -- a stand-in HIP runtime for CPU tests of the library's host code.
//
// The library's own objects (built by hipcc for gfx950) are linked against this
// file instead of libamdhip64, so the host side of every entry point runs in a
// process without a GPU.  Device memory is host memory carved from one large
// reserved region; copies are memcpy; kernels never run.  Every launch is
// recorded, and the pointer arguments of the pipeline's batched launches are
// checked against the live allocations: each required pointer must be non-null
// and the bytes the kernel may touch must lie inside one allocation.  Anything
// out of place is appended to a list of violations that the tests read.
//
// Test infrastructure only: nothing in the library or on the GPU path loads it.
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <string.h>

#include <cstdlib>
#include <cstring>
#include <sys/mman.h>

#include <cstdio>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../mtcheck_amd/csrc/args.hpp"

using namespace mtc;

namespace {

struct Alloc {
  uint64_t size;
  bool alive;
  bool device;
};

struct Node {
  const void* fn;
  dim3 grid, block;
  uint8_t arg[32];  // a Launch<A> (every kernel a graph may hold takes one)
};

struct Graph {
  std::vector<Node> nodes;
};

struct State {
  std::mutex mu;
  uint8_t* heap = nullptr;
  uint64_t heap_size = 1ull << 40;  // reserved, not committed
  uint64_t brk = 0;
  std::map<uintptr_t, Alloc> allocs;
  std::map<const void*, std::string> names;  // host stub -> device name
  std::map<const void*, std::pair<void*, uint64_t>> vars;  // host shadow -> (fake device address, size)
  std::vector<std::string> violations;
  std::vector<std::string> launches;  // kernel names in launch order
  std::vector<int> leaf_all_labels;   // per leaf member launched: its all_labels flag
  std::vector<int> label_resets;      // per k_assemble member launched: reset_labels is set
  std::vector<int> leaf_hash_later;   // per leaf member launched: it stores no digest
  std::vector<int> compact_hash;      // per k_compact member launched: it hashes the listed chunks
  // Words of a buffer known to hold zero, from its start (fills; k_compact clears
  // the leaf block counts it reads; a leaf launch makes its counts unknown).
  std::map<const void*, uint64_t> zero_upto;
  std::map<uintptr_t, uint64_t> host_allocs;  // pinned host blocks: start -> size
  hipStream_t capturing = nullptr;
  std::vector<Node> captured;
  uint64_t bytes_alive = 0;
};

State& S() {
  static State* s = new State();
  return *s;
}

thread_local dim3 g_grid, g_block;
thread_local size_t g_shmem = 0;
thread_local hipStream_t g_stream = nullptr;

void violation(const std::string& what) {
  S().violations.push_back(what);
}

// The allocation holding [p, p + bytes), or null.
const Alloc* find(const void* p, uint64_t bytes, uintptr_t* start_out = nullptr) {
  auto& st = S();
  const uintptr_t a = reinterpret_cast<uintptr_t>(p);
  auto it = st.allocs.upper_bound(a);
  if (it == st.allocs.begin()) return nullptr;
  --it;
  if (a < it->first || a >= it->first + std::max<uint64_t>(it->second.size, 1)) return nullptr;
  if (a + bytes > it->first + it->second.size) return nullptr;
  if (start_out) *start_out = it->first;
  return &it->second;
}

bool in_heap(const void* p) {
  const uintptr_t a = reinterpret_cast<uintptr_t>(p), h = reinterpret_cast<uintptr_t>(S().heap);
  return a >= h && a < h + S().heap_size;
}

// A pointer the kernel dereferences: non-null, inside a live allocation, with
// `bytes` (0: at least one byte) available from it.
bool need(const std::string& where, const char* field, const void* p, uint64_t bytes) {
  if (!p) {
    violation(where + ": " + field + " is null");
    return false;
  }
  const Alloc* a = find(p, std::max<uint64_t>(bytes, 1));
  if (!a) {
    const Alloc* b = find(p, 1);
    violation(where + ": " + field + (b ? " extends past its allocation (" + std::to_string(bytes) + " bytes)"
                                       : " points outside every allocation"));
    return false;
  }
  if (!a->alive) {
    violation(where + ": " + field + " points into freed memory");
    return false;
  }
  return true;
}

// A pointer the kernel may leave unused when it is null.
void maybe(const std::string& where, const char* field, const void* p, uint64_t bytes) {
  if (p) need(where, field, p, bytes);
}

struct LaunchRaw {
  const uint8_t* args;
  const uint32_t* prefix;
  const uint32_t* group;
  uint32_t count;
};

uint64_t tiles(uint64_t n) { return (n + 2047) / 2048; }

template <class A, class F>
void check_members(const std::string& name, const LaunchRaw& L, uint32_t grid, F&& fields) {
  bool ok = need(name, "Launch.args", L.args, sizeof(A) * (uint64_t)L.count);
  ok = need(name, "Launch.prefix", L.prefix, 4ull * (L.count + 1)) && ok;
  if (L.count > 1) ok = need(name, "Launch.group", L.group, 4ull * ((grid + 63) / 64)) && ok;
  if (!ok || L.count == 0) {
    if (ok) violation(name + ": a launch without members");
    return;
  }
  if (L.prefix[0] != 0 || L.prefix[L.count] != grid)
    violation(name + ": block offsets end at " + std::to_string(L.prefix[L.count]) + ", grid is " +
              std::to_string(grid));
  for (uint32_t m = 0; m < L.count; m++) {
    const uint32_t nblk = L.prefix[m + 1] - L.prefix[m];
    if (L.prefix[m + 1] < L.prefix[m]) violation(name + ": block offsets decrease");
    if (nblk == 0) continue;  // a member without blocks never reads its arguments
    A a;
    std::memcpy(&a, L.args + sizeof(A) * m, sizeof(A));
    fields(name + "[member " + std::to_string(m) + "]", a, nblk);
  }
  if (L.count > 1)
    for (uint64_t gi = 0; gi < (grid + 63) / 64; gi++) {
      const uint32_t m = L.group[gi];
      if (m >= L.count || L.prefix[m] > 64 * gi) violation(name + ": group table names a later member");
    }
}

const uint64_t kCtl = sizeof(Ctl);

void check_sweep(const std::string& w, const SweepArgs& a) {
  const uint64_t n = a.g.n;
  need(w, "leaf_dg", a.leaf_dg, 16 * n);
  need(w, "labels", a.labels, n);
  need(w, "leaf_slot", a.leaf_slot, 4 * n);
  const uint64_t tiles_out = 1ull << a.d_out;
  need(w, "out_lab", a.out_lab, tiles_out);
  need(w, "out_dg", a.out_dg, 16 * tiles_out);
  need(w, "out_slot", a.out_slot, 4 * tiles_out);
  if (!a.stage0) {
    const uint64_t in = 1ull << a.d_in;
    need(w, "in_lab", a.in_lab, in);
    need(w, "in_dg", a.in_dg, 16 * in);
    need(w, "in_slot", a.in_slot, 4 * in);
  }
  need(w, "tab", a.tab, 32ull * (a.mask + 2));
  need(w, "ctl", a.ctl, kCtl);
  need(w, "ctr", a.ctr, 8);
  need(w, "staging", a.staging, 8 * (2 * n - 1));
  need(w, "cnt_f", a.cnt_f, 4);
  need(w, "cnt_s", a.cnt_s, 4);
  need(w, "seg_base", a.seg_base, 4ull * (2 * a.g.D + 3));  // bases, then geometry
  // a stage-0 wavefront reads its tiles' flags with one lane per tile
  if (a.stage0 && (a.tiles_per_wave == 0 || a.tiles_per_wave > 64)) violation(w + ": tiles per wavefront out of 1..64");
  maybe(w, "dbg_tree", a.dbg_tree, 16 * (2 * n - 1));
  maybe(w, "dbg_labels", a.dbg_labels, 2 * n - 1);
  need(w, "tile_first", a.tile_first, 1);
  need(w, "tile_busy", a.tile_busy, 1);
  if (a.d_in <= a.d_out || a.d_in - a.d_out > 8) violation(w + ": stage depths out of range");
}

// A pinned host block (hipHostMalloc) holding [p, p + bytes).
void need_host(const std::string& where, const char* field, const void* p, uint64_t bytes) {
  auto& h = S().host_allocs;
  auto it = h.upper_bound(reinterpret_cast<uintptr_t>(p));
  if (!p || it == h.begin() || (--it, reinterpret_cast<uintptr_t>(p) + bytes > it->first + it->second))
    violation(where + ": " + field + " is not inside a pinned host block");
}

// A fill: its words hold the value (zero counts for the block-count tracking).
void note_fill(const FillSeg& f) {
  if (f.v != 0) return;
  uint64_t& z = S().zero_upto[f.p];
  z = std::max<uint64_t>(z, f.count);
}

void check_leaf(const std::string& w, const LeafArgs& a, uint32_t nblk, bool adds) {
  need(w, "data", a.data, a.len);
  need(w, "leaf_dg", a.leaf_dg, 16ull * a.n);
  need(w, "labels", a.labels, a.n);
  need(w, "list", a.list, 4ull * a.n);
  need(w, "blk_cnt", a.blk_cnt, 4ull * nblk);
  maybe(w, "list_dg", a.list_dg, 16ull * a.n);
  maybe(w, "stamp", a.stamp, 8);
  if ((uint64_t)nblk * a.tiles_per_block * 256 < a.n) violation(w + ": leaf blocks do not cover the chunks");
  uint64_t& z = S().zero_upto[a.blk_cnt];
  if (adds && z < nblk)
    violation(w + ": the leaf kernel adds to block counts not known to be zero (" + std::to_string(z) + " of " +
              std::to_string(nblk) + ")");
  z = 0;
  S().leaf_all_labels.push_back(a.all_labels);
  S().leaf_hash_later.push_back(a.hash_later);
}

void check_launch(const std::string& name, const uint8_t* arg0, uint32_t grid) {
  LaunchRaw L;
  std::memcpy(&L, arg0, sizeof(L));
  auto has = [&](const char* t) { return name.find(t) != std::string::npos; };
  if (has("7FillSeg")) {
    check_members<FillSeg>(name, L, grid, [](const std::string& w, const FillSeg& a, uint32_t) {
      if (need(w, "p", a.p, 4 * a.count)) note_fill(a);
    });
  } else if (has("7ScanSeg")) {
    check_members<ScanSeg>(name, L, grid, [](const std::string& w, const ScanSeg& a, uint32_t) {
      need(w, "in", a.in, 4ull * a.count_max);
      need(w, "out", a.out, 4ull * a.count_max);
      maybe(w, "count_dev", a.count_dev, 4);
      maybe(w, "total", a.total, 4);
      need(w, "sums", a.sums, 4 * tiles(a.count_max));
      maybe(w, "stamp", a.stamp, 8);
      maybe(w, "tile_lo", a.tile_lo, 4ull * (a.count_max / 256 + 1));  // region sizes: at most count_max chunks
    });
  } else if (has("8LeafArgs")) {
    const bool adds = has("7k_leaf1I");  // k_leaf1 adds to its block's count; k_leaf stores it
    check_members<LeafArgs>(name, L, grid, [adds](const std::string& w, const LeafArgs& a, uint32_t nblk) {
      check_leaf(w, a, nblk, adds);
    });
  } else if (has("11CompactArgs")) {
    check_members<CompactArgs>(name, L, grid, [](const std::string& w, const CompactArgs& a, uint32_t) {
      uint64_t& z = S().zero_upto[a.blk_cnt];
      z = std::max<uint64_t>(z, a.nblk);  // cleared after reading
      need(w, "list", a.list, 4);
      need(w, "blk_cnt", a.blk_cnt, 4ull * a.nblk);
      need(w, "blk_off", a.blk_off, 4ull * (a.nblk + 1));
      need(w, "dense", a.dense, 4);
      if (!a.list_dg != !a.dense_dg) violation(w + ": list_dg and dense_dg come together");
      maybe(w, "list_dg", a.list_dg, 16);
      maybe(w, "dense_dg", a.dense_dg, 16);
      if (a.hash_chunks) {  // digests from the chunk bytes, 16 bytes at a time
        need(w, "data", a.data, a.len);
        if (!a.dense_dg) violation(w + ": chunks hashed with nowhere to put the digests");
        if (a.C == 0 || a.C % 16 || reinterpret_cast<uintptr_t>(a.data) % 16) violation(w + ": chunks hashed with unaligned loads");
      }
      S().compact_hash.push_back(a.hash_chunks);
    });
  } else if (has("9DenseArgs")) {
    check_members<DenseArgs>(name, L, grid, [](const std::string& w, const DenseArgs& a, uint32_t) {
      need(w, "dense", a.dense, 4ull * a.n);
      need(w, "total", a.total, 4);
      need(w, "leaf_dg", a.leaf_dg, 16ull * a.n);
      need(w, "leaf_slot", a.leaf_slot, 4ull * a.n);
      need(w, "labels", a.labels, a.n);
      need(w, "dense_dg", a.dense_dg, 16ull * a.n);  // tree members' digests come only this way
      need(w, "pend", a.pend, 4ull * a.n);
      need(w, "pcnt", a.pcnt, (a.n + 63) / 64);
      need(w, "tab", a.tab, 32ull * (a.mask + 2));
      need(w, "ctl", a.ctl, kCtl);
      need(w, "ctr", a.ctr, 8);
      need(w, "tile_first", a.tile_first, 1);
      need(w, "tile_busy", a.tile_busy, 1);
    });
  } else if (has("k_single_leaf")) {
    check_members<SweepArgs>(name, L, grid, [](const std::string& w, const SweepArgs& a, uint32_t) {
      need(w, "labels", a.labels, 1);
      need(w, "leaf_slot", a.leaf_slot, 4);
      need(w, "staging", a.staging, 8);
      need(w, "cnt_f", a.cnt_f, 4);
      need(w, "cnt_s", a.cnt_s, 4);
    });
  } else if (has("9SweepTail")) {
    check_members<SweepTail>(name, L, grid, [](const std::string& w, const SweepTail& t, uint32_t nblk) {
      if (nblk != 1) violation(w + ": a sweep tail runs in one block");
      if (t.nst < 2 || t.nst > kTailStages) violation(w + ": sweep tail stage count out of 2..kTailStages");
      for (uint32_t i = 0; i < t.nst && i < kTailStages; i++) {
        const SweepArgs& a = t.st[i];
        check_sweep(w + "[stage " + std::to_string(i) + "]", a);
        const uint32_t tpw = a.stage0 ? a.tiles_per_wave : 1u;
        if (tpw && ((1ull << a.d_out) + tpw - 1) / tpw > kTailWaves) violation(w + ": a tail stage needs more wavefronts than its block has");
        if (i && a.d_in != t.st[i - 1].d_out) violation(w + ": tail stages do not follow each other");
      }
    });
  } else if (has("9SweepArgs")) {
    check_members<SweepArgs>(name, L, grid, [](const std::string& w, const SweepArgs& a, uint32_t) { check_sweep(w, a); });
  } else if (has("8RootArgs")) {
    check_members<RootArgs>(name, L, grid, [](const std::string& w, const RootArgs& a, uint32_t) {
      need(w, "staging", a.staging, 8);
      need(w, "cnt_f", a.cnt_f, 4ull * a.nseg);
      need(w, "cnt_s", a.cnt_s, 4ull * a.nseg);
      need(w, "off_f", a.off_f, 4ull * a.nseg);
      need(w, "off_s", a.off_s, 4ull * a.nseg);
      need(w, "seg_base", a.seg_base, 4ull * (2 * a.ndepth + 1));  // bases, then geometry
      need(w, "first_roots", a.first_roots, 4);
      need(w, "shift_tmp", a.shift_tmp, 8);
    });
  } else if (has("9ShiftArgs")) {
    check_members<ShiftArgs>(name, L, grid, [](const std::string& w, const ShiftArgs& a, uint32_t nblk) {
      maybe(w, "hist0", a.hist0, 4ull * a.nb0 * nblk);  // the first sort pass's block counts
      if (a.hist0 && (a.nb0 == 0 || a.nb0 > 256 || (a.nb0 & (a.nb0 - 1)))) violation(w + ": first-pass digits out of range");
      need(w, "shift_tmp", a.shift_tmp, 8);
      need(w, "S", a.S, 4);
      need(w, "leaf_slot", a.leaf_slot, 4ull * a.n);
      need(w, "tab", a.tab, 32ull * (a.mask + 2));
      need(w, "key", a.key, 4);
      need(w, "val", a.val, 8);
      need(w, "src_hist", a.src_hist, 4ull * a.nbins);
      need(w, "ctl", a.ctl, kCtl);
    });
  } else if (has("8SortArgs")) {
    check_members<SortArgs>(name, L, grid, [](const std::string& w, const SortArgs& a, uint32_t nblk) {
      need(w, "key_in", a.key_in, 4);
      need(w, "val_in", a.val_in, 8);
      need(w, "key_out", a.key_out, 4);
      need(w, "val_out", a.val_out, 8);
      need(w, "S", a.S, 4);
      need(w, "hist", a.hist, 4ull * a.nb * nblk);
    });
  } else if (has("10RegionArgs")) {
    check_members<RegionArgs>(name, L, grid, [](const std::string& w, const RegionArgs& a, uint32_t) {
      need(w, "first_roots", a.first_roots, 4);
      need(w, "F", a.F, 4);
      need(w, "sizes", a.sizes, 4ull * a.g.n);
      need(w, "pt.src_hist", a.pt.src_hist, 4ull * a.pt.nbins);
      need(w, "pt.ptable", a.pt.ptable, 8ull * a.pt.nbins);
      need(w, "pt.ctl", a.pt.ctl, kCtl);
    });
  } else if (has("13BasicDiffArgs")) {
    check_members<BasicDiffArgs>(name, L, grid, [](const std::string& w, const BasicDiffArgs& a, uint32_t) {
      need(w, "diff", a.diff, 32 + (uint64_t)a.n * (4 + a.C));
      need(w, "dense", a.dense, 4ull * a.n);
      need(w, "total", a.total, 4);
      need(w, "data", a.data, a.len);
      need(w, "ctl", a.ctl, kCtl);
      need(w, "ctr", a.ctr, 8);
      need_host(w, "out", a.out, kCtl);
    });
  } else if (has("8DiffArgs")) {
    check_members<DiffArgs>(name, L, grid, [](const std::string& w, const DiffArgs& a, uint32_t) {
      need(w, "diff", a.diff, 32 + (uint64_t)a.n * a.C);
      need(w, "first_roots", a.first_roots, 4);
      need(w, "ptable", a.ptable, 8);
      need(w, "shift_sorted", a.shift_sorted, 8);
      need(w, "region_off", a.region_off, 4ull * (a.n + 1));
      need(w, "tile_lo", a.tile_lo, 4ull * (a.n / 256 + 1));
      need(w, "total_chunks", a.total_chunks, 4);
      need(w, "data", a.data, a.len);
      need(w, "ctl", a.ctl, kCtl);
      need(w, "ctr", a.ctr, 8);
      need(w, "changed", a.changed, 4);
      need_host(w, "out", a.out, kCtl);
      need(w, "dense", a.dense, 4ull * a.n);
      maybe(w, "reset_labels", a.reset_labels, a.n);
      S().label_resets.push_back(a.reset_labels ? 1 : 0);
    });
  } else {
    violation(name + ": batched launch of an argument type this checker does not know");
  }
}

void check(const void* fn, dim3 grid, dim3 block, void** args) {
  auto& st = S();
  auto it = st.names.find(fn);
  const std::string name = it == st.names.end() ? std::string("<unregistered kernel>") : it->second;
  st.launches.push_back(name);
  if (grid.x == 0 || grid.y == 0 || grid.z == 0 || block.x == 0) violation(name + ": empty launch");
  if (block.x * block.y * block.z > 1024) violation(name + ": block too large");
  if (name.find("6Launch") != std::string::npos) {
    check_launch(name, static_cast<const uint8_t*>(args[0]), grid.x);
  } else if (name.find("k_leaf1_direct") != std::string::npos) {
    // The leaf kernel of a one-member step with its arguments by value: it copies
    // the step's argument arena from pinned host memory and does the step's fills.
    LeafArgs a;
    Prologue p;
    std::memcpy(&a, args[0], sizeof(a));
    std::memcpy(&p, args[1], sizeof(p));
    const uint64_t bytes = 16 * p.n16;
    if (bytes) {
      auto h = st.host_allocs.upper_bound(reinterpret_cast<uintptr_t>(p.src));
      if (h == st.host_allocs.begin() || (--h, reinterpret_cast<uintptr_t>(p.src) + bytes > h->first + h->second))
        violation(name + ": arena source is not inside a pinned host block");
      else if (need(name, "arena", p.dst, bytes))
        std::memcpy(p.dst, p.src, bytes);
    }
    if (p.nfill > kPrologueFills) violation(name + ": too many fills");
    // Its fills run beside its tiles, so they cannot clear the counts it adds to.
    check_leaf(name, a, grid.x, true);
    for (uint32_t f = 0; f < p.nfill && f < kPrologueFills; f++) {
      if (p.fill[f].p == a.blk_cnt) violation(name + ": fills the block counts it adds to");
      else if (need(name, "fill", p.fill[f].p, 4 * p.fill[f].count)) note_fill(p.fill[f]);
    }
  } else if (name.find("k_fill_u64") != std::string::npos) {
    uint64_t* p;
    uint64_t count;
    std::memcpy(&p, args[0], 8);
    std::memcpy(&count, args[1], 8);
    need(name, "p", p, 8 * count);
  }
}

}  // namespace

extern "C" {

// ---- test hooks --------------------------------------------------------------
int fakehip_violation_count() {
  std::lock_guard<std::mutex> g(S().mu);
  return (int)S().violations.size();
}
const char* fakehip_violation(int i) {
  std::lock_guard<std::mutex> g(S().mu);
  return (i >= 0 && i < (int)S().violations.size()) ? S().violations[i].c_str() : "";
}
int fakehip_launch_count() {
  std::lock_guard<std::mutex> g(S().mu);
  return (int)S().launches.size();
}
int fakehip_leaf_all_labels(int i) {  // -1 past the end
  std::lock_guard<std::mutex> g(S().mu);
  return (i >= 0 && i < (int)S().leaf_all_labels.size()) ? S().leaf_all_labels[i] : -1;
}
int fakehip_label_resets(int i) {
  std::lock_guard<std::mutex> g(S().mu);
  return (i >= 0 && i < (int)S().label_resets.size()) ? S().label_resets[i] : -1;
}
int fakehip_leaf_hash_later(int i) {
  std::lock_guard<std::mutex> g(S().mu);
  return (i >= 0 && i < (int)S().leaf_hash_later.size()) ? S().leaf_hash_later[i] : -1;
}
int fakehip_compact_hash(int i) {
  std::lock_guard<std::mutex> g(S().mu);
  return (i >= 0 && i < (int)S().compact_hash.size()) ? S().compact_hash[i] : -1;
}
const char* fakehip_launch_name(int i) {
  std::lock_guard<std::mutex> g(S().mu);
  return (i >= 0 && i < (int)S().launches.size()) ? S().launches[i].c_str() : "";
}
void fakehip_reset() {
  std::lock_guard<std::mutex> g(S().mu);
  S().violations.clear();
  S().launches.clear();
  S().leaf_all_labels.clear();
  S().label_resets.clear();
  S().leaf_hash_later.clear();
  S().compact_hash.clear();
}
uint64_t fakehip_bytes_alive() {
  std::lock_guard<std::mutex> g(S().mu);
  return S().bytes_alive;
}

// ---- registration --------------------------------------------------------------
void** __hipRegisterFatBinary(const void*) {
  static void* handle = nullptr;
  return &handle;
}
void __hipUnregisterFatBinary(void**) {}
void __hipRegisterFunction(void**, const void* hostFunction, char*, const char* deviceName, unsigned int, uint3*,
                           uint3*, dim3*, dim3*, int*) {
  std::lock_guard<std::mutex> g(S().mu);
  S().names[hostFunction] = deviceName ? deviceName : "";
}
void __hipRegisterVar(void**, void* var, char*, char*, int, size_t size, int, int) {
  // The device copy of a variable: an allocation of its own in the fake heap.
  void* p = nullptr;
  hipMalloc(&p, size);
  std::lock_guard<std::mutex> g(S().mu);
  S().vars[var] = {p, size};
}
void __hipRegisterManagedVar(void*, void**, void*, const char*, size_t, unsigned) {}
void __hipRegisterSurface(void*, void*, char*, const char*, int, int) {}
void __hipRegisterTexture(void*, void*, char*, const char*, int, int, int) {}

hipError_t __hipPushCallConfiguration(dim3 gridDim, dim3 blockDim, size_t sharedMem, hipStream_t stream) {
  g_grid = gridDim;
  g_block = blockDim;
  g_shmem = sharedMem;
  g_stream = stream;
  return hipSuccess;
}
hipError_t __hipPopCallConfiguration(dim3* gridDim, dim3* blockDim, size_t* sharedMem, hipStream_t* stream) {
  *gridDim = g_grid;
  *blockDim = g_block;
  *sharedMem = g_shmem;
  *stream = g_stream;
  return hipSuccess;
}

hipError_t hipLaunchKernel(const void* fn, dim3 grid, dim3 block, void** args, size_t, hipStream_t stream) {
  std::lock_guard<std::mutex> g(S().mu);
  auto& st = S();
  if (st.capturing && stream == st.capturing) {
    Node nd;
    nd.fn = fn;
    nd.grid = grid;
    nd.block = block;
    auto it = st.names.find(fn);
    if (it == st.names.end() || it->second.find("6Launch") == std::string::npos) {
      violation("a kernel without a Launch argument was captured into a graph");
      return hipSuccess;
    }
    std::memcpy(nd.arg, args[0], sizeof(nd.arg));
    st.captured.push_back(nd);
    return hipSuccess;
  }
  check(fn, grid, block, args);
  return hipSuccess;
}

// ---- memory --------------------------------------------------------------------
hipError_t hipMalloc(void** ptr, size_t size) {
  std::lock_guard<std::mutex> g(S().mu);
  auto& st = S();
  if (!st.heap) {
    void* h = mmap(nullptr, st.heap_size, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE, -1, 0);
    if (h == MAP_FAILED) return hipErrorOutOfMemory;
    st.heap = static_cast<uint8_t*>(h);
  }
  const uint64_t sz = size ? size : 1;
  const uint64_t off = (st.brk + 4095) & ~4095ull;  // page aligned, one guard page between allocations
  if (off + sz + 4096 > st.heap_size) return hipErrorOutOfMemory;
  st.brk = off + sz + 4096;
  *ptr = st.heap + off;
  st.allocs[reinterpret_cast<uintptr_t>(*ptr)] = Alloc{sz, true, true};
  st.bytes_alive += sz;
  return hipSuccess;
}
hipError_t hipFree(void* p) {
  if (!p) return hipSuccess;
  std::lock_guard<std::mutex> g(S().mu);
  auto& st = S();
  auto it = st.allocs.find(reinterpret_cast<uintptr_t>(p));
  if (it == st.allocs.end() || !it->second.alive) {
    violation("hipFree of a pointer that is not a live allocation");
    return hipErrorInvalidValue;
  }
  it->second.alive = false;
  st.bytes_alive -= it->second.size;
  madvise(p, (it->second.size + 4095) & ~4095ull, MADV_DONTNEED);
  return hipSuccess;
}
hipError_t hipHostMalloc(void** ptr, size_t size, unsigned int) {
  *ptr = std::calloc(1, size ? size : 1);
  if (!*ptr) return hipErrorOutOfMemory;
  std::lock_guard<std::mutex> g(S().mu);
  S().host_allocs[reinterpret_cast<uintptr_t>(*ptr)] = size;
  return hipSuccess;
}
hipError_t hipHostFree(void* p) {
  {
    std::lock_guard<std::mutex> g(S().mu);
    S().host_allocs.erase(reinterpret_cast<uintptr_t>(p));
  }
  std::free(p);
  return hipSuccess;
}

static hipError_t copy(void* dst, const void* src, size_t n) {
  {
    std::lock_guard<std::mutex> g(S().mu);
    if (in_heap(dst) && !find(dst, n)) violation("copy writes past a device allocation");
    if (in_heap(src) && !find(src, n)) violation("copy reads past a device allocation");
  }
  if (n) std::memmove(dst, src, n);
  return hipSuccess;
}
hipError_t hipMemcpyAsync(void* dst, const void* src, size_t n, hipMemcpyKind, hipStream_t) {
  return copy(dst, src, n);
}
hipError_t hipMemcpy(void* dst, const void* src, size_t n, hipMemcpyKind) { return copy(dst, src, n); }
hipError_t hipMemsetAsync(void* dst, int v, size_t n, hipStream_t) {
  {
    std::lock_guard<std::mutex> g(S().mu);
    if (!find(dst, n)) violation("memset past a device allocation");
  }
  std::memset(dst, v, n);
  return hipSuccess;
}
hipError_t hipGetSymbolAddress(void** devPtr, const void* symbol) {
  std::lock_guard<std::mutex> g(S().mu);
  auto it = S().vars.find(symbol);
  if (it == S().vars.end()) return hipErrorInvalidSymbol;
  *devPtr = it->second.first;
  return hipSuccess;
}
hipError_t hipPointerGetAttributes(hipPointerAttribute_t* attr, const void*) {
  std::memset(attr, 0, sizeof(*attr));
  attr->type = hipMemoryTypeDevice;
  attr->device = 0;
  return hipSuccess;
}

// ---- devices, streams, events, graphs ------------------------------------------
hipError_t hipGetDeviceCount(int* c) {
  *c = 1;
  return hipSuccess;
}
hipError_t hipGetDevice(int* d) {
  *d = 0;
  return hipSuccess;
}
hipError_t hipSetDevice(int d) { return d == 0 ? hipSuccess : hipErrorInvalidDevice; }
hipError_t hipDeviceSynchronize() { return hipSuccess; }
hipError_t hipDeviceGetAttribute(int* v, hipDeviceAttribute_t attr, int) {
  switch (attr) {
    case hipDeviceAttributeMultiprocessorCount: *v = 256; break;
    case hipDeviceAttributeWallClockRate: *v = 100000; break;
    default: *v = 0; break;
  }
  return hipSuccess;
}

hipError_t hipGetLastError() { return hipSuccess; }
const char* hipGetErrorString(hipError_t) { return "fake HIP error"; }

hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned int) {
  *s = reinterpret_cast<hipStream_t>(new uint64_t(0));
  return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t s) {
  delete reinterpret_cast<uint64_t*>(s);
  return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned int) { return hipSuccess; }
hipError_t hipEventCreate(hipEvent_t* e) {
  *e = reinterpret_cast<hipEvent_t>(new uint64_t(0));
  return hipSuccess;
}
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned int) { return hipEventCreate(e); }
hipError_t hipEventDestroy(hipEvent_t e) {
  delete reinterpret_cast<uint64_t*>(e);
  return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
hipError_t hipEventElapsedTime(float* ms, hipEvent_t, hipEvent_t) {
  *ms = 0.0f;
  return hipSuccess;
}

hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode) {
  std::lock_guard<std::mutex> g(S().mu);
  S().capturing = s;
  S().captured.clear();
  return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t* graph) {
  std::lock_guard<std::mutex> g(S().mu);
  Graph* gr = new Graph{S().captured};
  S().captured.clear();
  S().capturing = nullptr;
  *graph = reinterpret_cast<hipGraph_t>(gr);
  return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t* exec, hipGraph_t graph, hipGraphNode_t*, char*, size_t) {
  *exec = reinterpret_cast<hipGraphExec_t>(new Graph(*reinterpret_cast<Graph*>(graph)));
  return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t graph) {
  delete reinterpret_cast<Graph*>(graph);
  return hipSuccess;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t exec) {
  delete reinterpret_cast<Graph*>(exec);
  return hipSuccess;
}
// A replay checks every captured launch against the arena as it is now.
hipError_t hipGraphLaunch(hipGraphExec_t exec, hipStream_t) {
  std::lock_guard<std::mutex> g(S().mu);
  for (Node& nd : reinterpret_cast<Graph*>(exec)->nodes) {
    void* args[1] = {nd.arg};
    check(nd.fn, nd.grid, nd.block, args);
  }
  return hipSuccess;
}

}  // extern "C"
