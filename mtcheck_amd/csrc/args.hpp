// args.hpp -- argument blocks of the checkpoint pipeline's batched launches.
//
// Every pipeline kernel takes a Launch<A>: one argument block A per member (a
// deduplicator in the step) and the block offsets of each member's share of the
// grid.  The host writes the blocks into one arena that is uploaded per step.
#pragma once
#include "internal.hpp"

namespace mtc {

template <class A>
struct Launch {
  const A* args;           // one per member
  const uint32_t* prefix;  // count + 1 block offsets
  const uint32_t* group;   // member owning block 64*i (only when count > 1)
  uint32_t count;
};

struct Where {
  uint32_t m, blk, nblk;  // member, block within the member's share, share size
};

// Pointers inside argument blocks are loaded from memory, so the compiler cannot
// tell that they address global memory and would emit generic (flat) accesses.
// A round trip through the global address space lets it use global instructions.
template <class T>
__device__ __forceinline__ T* gl(T* p) {
  return (T*)(__attribute__((address_space(1))) T*)(uintptr_t)p;
}
#define MTC_GLOBAL(f) c.f = gl(c.f)

struct FillSeg {
  uint32_t* p;
  uint64_t count;  // 32-bit words
  uint32_t v;
  __device__ FillSeg global() const {
    FillSeg c = *this;
    MTC_GLOBAL(p);
    return c;
  }
};

struct ScanSeg {
  const uint32_t* in;
  uint32_t* out;          // may equal in
  const uint32_t* count_dev;
  uint32_t count_max;
  uint32_t* total;        // optional: sum of all elements
  uint32_t* sums;         // ceil(count_max / kScanTile) words
  unsigned long long* stamp;  // optional: the wall clock when the segment's first block starts
  // Optional, for a scan of region sizes: tile_lo[b] = the element whose range of
  // the running total holds position 256 b (the first region of diff tile b).
  uint32_t* tile_lo;
  __device__ ScanSeg global() const {
    ScanSeg c = *this;
    MTC_GLOBAL(in); MTC_GLOBAL(out); MTC_GLOBAL(count_dev); MTC_GLOBAL(total); MTC_GLOBAL(sums); MTC_GLOBAL(stamp);
    MTC_GLOBAL(tile_lo);
    return c;
  }
};

struct LeafArgs {
  const uint8_t* data;
  uint64_t len;
  uint32_t n;
  uint32_t C;
  Digest* leaf_dg;
  uint8_t* labels;
  int baseline;       // 1: previous digests are zero (a reset); 2: every chunk counts as changed
  uint32_t* list;     // chunks to insert, packed per block
  uint32_t* blk_cnt;  // their count per block
  uint64_t tiles_per_block;
  // 1: label every chunk.  0: the labels already hold L_FIXED for every chunk (the
  // previous step reset the ones it listed) and no chunk is labelled here: k_insert
  // labels the listed ones
  int all_labels;
  // Tree members: each listed chunk's digest is stored beside its list entry and
  // k_insert moves it to leaf_dg.  Null: leaf_dg is written here.
  Digest* list_dg;
  unsigned long long* stamp;  // the wall clock when the member's first block starts
  // 1 (tree members, k_leaf1): no digest is stored here; k_compact hashes the
  // listed chunks again (they are few when few chunks change)
  int hash_later;
  __device__ LeafArgs global() const {
    LeafArgs c = *this;
    MTC_GLOBAL(data); MTC_GLOBAL(leaf_dg); MTC_GLOBAL(labels); MTC_GLOBAL(list); MTC_GLOBAL(blk_cnt); MTC_GLOBAL(list_dg);
    MTC_GLOBAL(stamp);
    return c;
  }
};

// What a direct leaf launch does before its tile: the arena copy (16-byte words,
// from pinned host memory) and the step's fills.
constexpr uint32_t kPrologueFills = 12;
struct Prologue {
  const uint4* src;
  uint4* dst;
  uint64_t n16;
  FillSeg fill[kPrologueFills];
  uint32_t nfill;
};

struct CompactArgs {
  const uint32_t* list;
  uint32_t* blk_cnt;  // read, then cleared for the next step's leaf kernel
  const uint32_t* blk_off;
  uint64_t tiles_per_block;
  uint32_t* dense;
  uint32_t nblk;            // leaf blocks
  const Digest* list_dg;    // digests beside the list entries (tree members), or null
  Digest* dense_dg;         // their dense copy
  // hash_chunks: dense_dg from the chunk bytes (Murmur3; 16-byte aligned data, C a
  // multiple of 16) instead of list_dg
  const uint8_t* data;
  uint64_t len;
  uint32_t C;
  int hash_chunks;
  __device__ CompactArgs global() const {
    CompactArgs c = *this;
    MTC_GLOBAL(list); MTC_GLOBAL(blk_cnt); MTC_GLOBAL(blk_off); MTC_GLOBAL(dense); MTC_GLOBAL(list_dg); MTC_GLOBAL(dense_dg);
    MTC_GLOBAL(data);
    return c;
  }
};

struct DenseArgs {
  const uint32_t* dense;  // chunks to insert, ascending within each wavefront's run
  const uint32_t* total;  // their number
  uint32_t n;
  Digest* leaf_dg;         // written here from dense_dg
  uint32_t* leaf_slot;
  uint8_t* labels;
  Slot* tab;
  uint64_t mask;
  uint32_t cur;
  Ctl* ctl;
  unsigned long long* ctr;
  Geo g;
  uint32_t R0;          // levels of the first sweep stage
  uint8_t* tile_first;  // per stage-0 tile: holds a first-occurrence leaf
  uint8_t* tile_busy;   // per stage-0 tile: holds a leaf that is not a fixed duplicate
  const Digest* dense_dg;  // the listed chunks' digests, in dense order
  int leaf_labels;         // the leaf kernel labelled the chunks (otherwise they count as pending)
  uint32_t* pend;          // k_insert lists the chunks k_fixmin and k_resolve visit, each group of
                           // 64 dense entries at the front of its own 64 words ...
  uint8_t* pcnt;           // ... and counts them per group (ceil(n / 64) bytes)
  __device__ DenseArgs global() const {
    DenseArgs c = *this;
    MTC_GLOBAL(dense); MTC_GLOBAL(total); MTC_GLOBAL(leaf_dg); MTC_GLOBAL(leaf_slot); MTC_GLOBAL(labels); MTC_GLOBAL(tab); MTC_GLOBAL(ctl); MTC_GLOBAL(ctr); MTC_GLOBAL(tile_first); MTC_GLOBAL(tile_busy); MTC_GLOBAL(dense_dg); MTC_GLOBAL(pend); MTC_GLOBAL(pcnt);
    return c;
  }
};

struct SweepArgs {
  Geo g;
  uint32_t cur;
  uint32_t d_in, d_out;
  int stage0;
  const Digest* leaf_dg;
  uint8_t* labels;
  const uint32_t* leaf_slot;
  const uint8_t* in_lab;
  const Digest* in_dg;
  const uint32_t* in_slot;
  uint8_t* out_lab;
  Digest* out_dg;
  uint32_t* out_slot;
  Slot* tab;
  uint64_t mask;
  Ctl* ctl;
  unsigned long long* ctr;
  uint64_t* staging;
  uint32_t* cnt_f;
  uint32_t* cnt_s;
  const uint32_t* seg_base;
  Digest* dbg_tree;
  uint8_t* dbg_labels;
  const uint8_t* tile_first;  // stage 0: tile holds a first-occurrence leaf
  const uint8_t* tile_busy;   // stage 0: tile holds a leaf that is not a fixed duplicate
  uint32_t tiles_per_wave;     // stage 0: consecutive tiles one wavefront walks (<= 64)
  int compat;                  // node 0 is a root only when it is a first occurrence
  __device__ SweepArgs global() const {
    SweepArgs c = *this;
    MTC_GLOBAL(leaf_dg); MTC_GLOBAL(labels); MTC_GLOBAL(leaf_slot); MTC_GLOBAL(in_lab); MTC_GLOBAL(in_dg); MTC_GLOBAL(in_slot); MTC_GLOBAL(out_lab); MTC_GLOBAL(out_dg); MTC_GLOBAL(out_slot); MTC_GLOBAL(tab); MTC_GLOBAL(ctl); MTC_GLOBAL(ctr); MTC_GLOBAL(staging); MTC_GLOBAL(cnt_f); MTC_GLOBAL(cnt_s); MTC_GLOBAL(seg_base); MTC_GLOBAL(dbg_tree); MTC_GLOBAL(dbg_labels); MTC_GLOBAL(tile_first); MTC_GLOBAL(tile_busy);
    return c;
  }
};

// The last sweep stages of a member, each small enough for one block of
// kTailWaves wavefronts, run by one block in order (a barrier between stages).
constexpr uint32_t kTailStages = 8, kTailWaves = 8;
struct SweepTail {
  SweepArgs st[kTailStages];
  uint32_t nst;
};

struct RootArgs {
  const uint64_t* staging;
  const uint32_t *cnt_f, *cnt_s, *off_f, *off_s, *seg_base;
  uint32_t ndepth, nseg;
  uint32_t* first_roots;
  uint64_t* shift_tmp;
  __device__ RootArgs global() const {
    RootArgs c = *this;
    MTC_GLOBAL(staging); MTC_GLOBAL(cnt_f); MTC_GLOBAL(cnt_s); MTC_GLOBAL(off_f); MTC_GLOBAL(off_s); MTC_GLOBAL(seg_base); MTC_GLOBAL(first_roots); MTC_GLOBAL(shift_tmp);
    return c;
  }
};

struct ShiftArgs {
  const uint64_t* shift_tmp;
  const uint32_t* S;
  const uint32_t* leaf_slot;
  uint32_t n;
  const Slot* tab;
  uint64_t mask;
  uint32_t ref, cur;
  uint32_t* key;
  uint64_t* val;
  uint32_t* src_hist;
  uint32_t nbins;
  Ctl* ctl;
  // When set: the first sort pass's per-block digit counts (nb0 digit values,
  // digit-major over the grid), written here instead of by k_sort_count; the
  // entries are then visited in the sort's block ranges.
  uint32_t* hist0;
  uint32_t nb0;
  __device__ ShiftArgs global() const {
    ShiftArgs c = *this;
    MTC_GLOBAL(shift_tmp); MTC_GLOBAL(S); MTC_GLOBAL(leaf_slot); MTC_GLOBAL(tab); MTC_GLOBAL(key); MTC_GLOBAL(val); MTC_GLOBAL(src_hist); MTC_GLOBAL(ctl);
    MTC_GLOBAL(hist0);
    return c;
  }
};

struct SortArgs {
  const uint32_t* key_in;
  const uint64_t* val_in;
  uint32_t* key_out;
  uint64_t* val_out;
  const uint32_t* S;
  uint32_t shift, nb;  // digit shift and number of digit values (a power of two <= 256)
  uint32_t* hist;      // nb * G counts, then their exclusive scan
  __device__ SortArgs global() const {
    SortArgs c = *this;
    MTC_GLOBAL(key_in); MTC_GLOBAL(val_in); MTC_GLOBAL(key_out); MTC_GLOBAL(val_out); MTC_GLOBAL(S); MTC_GLOBAL(hist);
    return c;
  }
};

struct PtableArgs {
  const uint32_t* src_hist;
  uint32_t nbins, ref;
  uint32_t* ptable;
  Ctl* ctl;
  __device__ PtableArgs global() const {
    PtableArgs c = *this;
    MTC_GLOBAL(src_hist); MTC_GLOBAL(ptable); MTC_GLOBAL(ctl);
    return c;
  }
};

struct RegionArgs {
  const uint32_t* first_roots;
  const uint32_t* F;
  Geo g;
  uint32_t* sizes;
  int compat;     // the reference's leaf ranges
  PtableArgs pt;  // the (chkpt, count) table, built by block 0
  __device__ RegionArgs global() const {
    RegionArgs c = *this;
    MTC_GLOBAL(first_roots); MTC_GLOBAL(F); MTC_GLOBAL(sizes);
    c.pt = pt.global();
    return c;
  }
};

struct DiffArgs {
  uint8_t* diff;
  const uint32_t* first_roots;
  const uint32_t* ptable;
  const uint64_t* shift_sorted;  // (prev << 32 | node)
  const uint32_t* region_off;
  const uint32_t* tile_lo;  // per 256 diff chunks: the region holding the first
  const uint32_t* total_chunks;
  const uint8_t* data;
  uint64_t len;
  uint32_t n, C;
  uint32_t ref_id, chkpt_id;
  Geo g;
  Ctl* ctl;
  int aligned;
  int compat;                     // the reference's leaf ranges
  const unsigned long long* ctr;  // striped counters
  const uint32_t* changed;        // chunks inserted by the leaf pass
  Ctl* out;                       // the member's entry of the step's result array
  const uint32_t* dense;          // the listed chunks (*changed of them)
  uint8_t* reset_labels;          // when set: their labels go back to L_FIXED
  __device__ DiffArgs global() const {
    DiffArgs c = *this;
    MTC_GLOBAL(diff); MTC_GLOBAL(first_roots); MTC_GLOBAL(ptable); MTC_GLOBAL(shift_sorted); MTC_GLOBAL(region_off); MTC_GLOBAL(tile_lo); MTC_GLOBAL(total_chunks); MTC_GLOBAL(data); MTC_GLOBAL(ctl);
    MTC_GLOBAL(ctr); MTC_GLOBAL(changed); MTC_GLOBAL(out); MTC_GLOBAL(dense); MTC_GLOBAL(reset_labels);
    return c;
  }
};

struct BasicDiffArgs {
  uint8_t* diff;
  const uint32_t* dense;  // changed chunks, ascending
  const uint32_t* total;  // their number
  const uint8_t* data;
  uint64_t len;
  uint32_t n, C;
  uint32_t ref_id, chkpt_id;
  Ctl* ctl;
  int aligned;
  const unsigned long long* ctr;
  Ctl* out;
  __device__ BasicDiffArgs global() const {
    BasicDiffArgs c = *this;
    MTC_GLOBAL(diff); MTC_GLOBAL(dense); MTC_GLOBAL(total); MTC_GLOBAL(data); MTC_GLOBAL(ctl); MTC_GLOBAL(ctr); MTC_GLOBAL(out);
    return c;
  }
};

}  // namespace mtc
