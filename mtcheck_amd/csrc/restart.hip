// restart.hip -- rebuild checkpoint k from the diff chain ref_id..k.
//
// Semantics: src/tree_approach.cpp:712-1128.  A chunk takes its bytes from the
// newest diff of the chain that covers its leaf, either from that diff's own
// first-occurrence data or, for a shifted-duplicate entry (node, prev) with source
// checkpoint s, from the first-occurrence region of diff s that contains prev.
// Instead of the reference's per-diff pass over every chunk, the chain is resolved
// once into a per-chunk (diff, data slot) cover array, then one scatter copy runs.
#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>

#include "internal.hpp"

namespace mtc {

constexpr uint32_t kRThreads = 256;
constexpr uint64_t kSlotBits = 40;
constexpr uint64_t kSlotMask = (1ull << kSlotBits) - 1;
constexpr uint64_t kNoCover = ~0ull;

struct ChainDiff {
  const uint8_t* base;     // device
  uint64_t size;
  mtc_header h;
  uint64_t off_first, off_table, off_shift, off_data;
  uint64_t data_chunks;    // whole chunks the data section can hold
  uint64_t region_base;    // first entry of this diff in the flat region-offset array
};

static uint32_t rgrid(uint64_t items, uint32_t cap = 1u << 20) {
  uint64_t g = (items + kRThreads - 1) / kRThreads;
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(g, cap));
}

__device__ __forceinline__ uint32_t ld32(const uint8_t* p) {
  return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}

// Hash set (diff position, root node) -> region index, for prev lookups.
__device__ __forceinline__ uint64_t rs_home(uint64_t key, uint64_t mask) {
  key ^= key >> 33;
  key *= 0xff51afd7ed558ccdULL;
  key ^= key >> 33;
  return key & mask;
}

// One diff's first-occurrence roots: the region size of each, and its entry in
// the (diff position, root node) set that resolves shifted entries' sources.
// idb: added to every stored id (0 for tree diffs, whose ids are node ids; n - 1
// for basic diffs, whose ids are chunk indices: leaf n - 1 + c holds chunk c).
__global__ void k_r_roots(const uint8_t* first, uint32_t F, Geo g, uint32_t* sizes, uint32_t pos, uint64_t* keys,
                          uint32_t* vals, uint64_t mask, uint32_t* err, uint32_t idb) {
  const uint32_t i = blockIdx.x * kRThreads + threadIdx.x;
  if (i >= F) return;
  const uint64_t node = (uint64_t)ld32(first + 4ull * i) + idb;
  if (node >= 2ull * g.n - 1) {
    atomicOr(err, ERR_CORRUPT);
    sizes[i] = 0;
    return;
  }
  sizes[i] = leaf_count(g, (uint32_t)node);
  const uint64_t key = ((uint64_t)pos << 32) | node;
  uint64_t s = rs_home(key, mask);
  for (uint64_t probes = 0; probes <= mask; probes++) {
    const uint64_t old = atomicCAS((unsigned long long*)&keys[s], (unsigned long long)kNoCover, (unsigned long long)key);
    if (old == kNoCover) {
      vals[s] = i;
      return;
    }
    if (old == key) {  // the same root twice in one diff
      atomicOr(err, ERR_CORRUPT);
      return;
    }
    s = (s + 1) & mask;
  }
  atomicOr(err, ERR_TABLE_FULL);
}

__device__ __forceinline__ uint32_t rs_find(const uint64_t* keys, const uint32_t* vals, uint64_t mask, uint64_t key) {
  uint64_t s = rs_home(key, mask);
  for (uint64_t probes = 0; probes <= mask; probes++) {
    const uint64_t k = keys[s];
    if (k == key) return vals[s];
    if (k == kNoCover) return ~0u;
    s = (s + 1) & mask;
  }
  return ~0u;
}

__global__ void k_r_fill64(uint64_t* p, uint64_t count, uint64_t v) {
  for (uint64_t i = (uint64_t)blockIdx.x * kRThreads + threadIdx.x; i < count; i += (uint64_t)gridDim.x * kRThreads)
    p[i] = v;
}

__global__ void k_r_fill32(uint32_t* p, uint64_t count, uint32_t v) {
  for (uint64_t i = (uint64_t)blockIdx.x * kRThreads + threadIdx.x; i < count; i += (uint64_t)gridDim.x * kRThreads)
    p[i] = v;
}

// Chunks of one diff's first-occurrence regions: cover[chunk] = (pos, data slot).
__global__ void k_r_cover_first(const uint8_t* first, uint32_t F, const uint32_t* region_off, uint64_t total, Geo g,
                                uint32_t pos, uint64_t* cover, uint32_t idb) {
  for (uint64_t j = (uint64_t)blockIdx.x * kRThreads + threadIdx.x; j < total; j += (uint64_t)gridDim.x * kRThreads) {
    uint32_t lo = 0, hi = F;
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (region_off[mid] <= j) lo = mid;
      else hi = mid;
    }
    uint32_t st[2], ct[2];
    const int r = leaf_runs(g, ld32(first + 4ull * lo) + idb, st, ct);
    const uint32_t chunk = run_at(r, st, ct, (uint32_t)(j - region_off[lo]));
    cover[chunk] = ((uint64_t)pos << kSlotBits) | j;
  }
}

struct RShiftArgs {
  const uint8_t* shift;   // S (node, prev) pairs of this diff
  uint32_t S;
  const uint32_t* cum;    // inclusive running count by checkpoint id, ref..chkpt
  uint32_t ncum;
  uint32_t ref;
  uint32_t pos;
  Geo g;
  const uint64_t* set_keys;
  const uint32_t* set_vals;
  uint64_t set_mask;
  const uint64_t* first_ptr;   // per chain position: device address of its F roots
  const uint64_t* region_base; // per chain position: first entry in region_off_all
  const uint32_t* region_off_all;
  uint64_t* cover;
  uint32_t* err;
};

// Shifted-duplicate entries of one diff: each leaf of `node` takes the data slot
// of the matching leaf of `prev` inside the source diff's region
// (src/tree_approach.cpp:887-925, 1058-1118).
__global__ void k_r_cover_shift(RShiftArgs a) {
  const uint32_t i = blockIdx.x * kRThreads + threadIdx.x;
  if (i >= a.S) return;
  const uint32_t node = ld32(a.shift + 8ull * i), prev = ld32(a.shift + 8ull * i + 4);
  const uint64_t nodes = 2ull * a.g.n - 1;
  if (node >= nodes || prev >= nodes) {
    atomicOr(a.err, ERR_CORRUPT);
    return;
  }
  uint32_t lo = 0, hi = a.ncum;  // smallest checkpoint whose running count exceeds i
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (i < a.cum[mid]) hi = mid;
    else lo = mid + 1;
  }
  if (lo >= a.ncum || lo > a.pos) {
    atomicOr(a.err, ERR_CORRUPT);
    return;
  }
  const uint32_t spos = lo;  // chain position of the source checkpoint
  uint32_t anc = prev, region = ~0u;
  while (true) {
    region = rs_find(a.set_keys, a.set_vals, a.set_mask, ((uint64_t)spos << 32) | anc);
    if (region != ~0u || anc == 0) break;
    anc = (anc - 1) >> 1;
  }
  if (region == ~0u) {
    atomicOr(a.err, ERR_CORRUPT);
    return;
  }
  const uint8_t* sfirst = reinterpret_cast<const uint8_t*>(a.first_ptr[spos]);
  const uint32_t root = ld32(sfirst + 4ull * region);
  const uint32_t rbase = a.region_off_all[a.region_base[spos] + region];
  uint32_t rs[2], rc[2], ns[2], nc[2], ps[2], pc[2];
  const int rr = leaf_runs(a.g, root, rs, rc);
  const int nr = leaf_runs(a.g, node, ns, nc);
  const int pr = leaf_runs(a.g, prev, ps, pc);
  const uint32_t cnt = nc[0] + (nr == 2 ? nc[1] : 0);
  if (cnt != pc[0] + (pr == 2 ? pc[1] : 0)) {
    atomicOr(a.err, ERR_CORRUPT);
    return;
  }
  for (uint32_t j = 0; j < cnt; j++) {
    const uint32_t dst_chunk = run_at(nr, ns, nc, j);
    const uint32_t src_chunk = run_at(pr, ps, pc, j);
    const uint64_t slot = rbase + run_rank(rr, rs, rc, src_chunk);
    a.cover[dst_chunk] = ((uint64_t)spos << kSlotBits) | slot;
  }
}


// One scatter copy: out[chunk] = data section of diff pos, slot.
__global__ __launch_bounds__(kRThreads) void k_r_copy(const uint64_t* cover, uint32_t n, uint32_t C, uint64_t datalen,
                                                       const uint64_t* data_ptr, const uint64_t* data_chunks,
                                                       uint8_t* out, int vec, uint32_t* err) {
  __shared__ uint64_t s_src[kRThreads];
  const uint32_t U = vec ? C / 16 : C;
  for (uint64_t tile = blockIdx.x; tile * kRThreads < n; tile += gridDim.x) {
    const uint64_t c0 = tile * kRThreads;
    const uint64_t c = c0 + threadIdx.x;
    if (c < n) {
      const uint64_t e = cover[c];
      uint64_t src = 0;
      if (e == kNoCover) {
        atomicOr(err, ERR_CORRUPT);
      } else {
        const uint32_t pos = (uint32_t)(e >> kSlotBits);
        const uint64_t slot = e & kSlotMask;
        if (slot >= data_chunks[pos]) atomicOr(err, ERR_CORRUPT);
        else src = data_ptr[pos] + slot * C;
      }
      s_src[threadIdx.x] = src;
    }
    __syncthreads();
    const uint64_t cnt = ((uint64_t)n - c0 < kRThreads) ? (uint64_t)n - c0 : kRThreads;
    for (uint64_t u = threadIdx.x; u < cnt * U; u += kRThreads) {
      const uint32_t jj = (uint32_t)(u / U), w = (uint32_t)(u % U);
      const uint64_t src = s_src[jj];
      if (!src) continue;
      const uint64_t doff = (c0 + jj) * C;
      const uint8_t* sp = reinterpret_cast<const uint8_t*>(src);
      if (vec) {
        const uint64_t b0 = doff + 16ull * w;
        if (b0 + 16 <= datalen) {
          st_stream16(out + b0, ld_stream16_a4(sp + 16ull * w));
        } else {
          for (uint32_t b = 0; b < 16 && b0 + b < datalen; b++) out[b0 + b] = sp[16ull * w + b];
        }
      } else if (doff + w < datalen) {
        out[doff + w] = sp[w];
      }
    }
    __syncthreads();
  }
}

struct RestartScratch {
  hipEvent_t ev[3] = {nullptr, nullptr, nullptr};  // start, chain uploaded, done
  Workspace ws;
  DevBuf<uint32_t> err, region_off, set_vals;
  DevBuf<uint64_t> set_keys, cover, meta;
  uint8_t* stage = nullptr;  // pinned: per-position tables for the kernels
  uint64_t stage_cap = 0;
  uint8_t* hstage = nullptr;  // pinned: headers and prior-checkpoint tables of device chains
  uint64_t hstage_cap = 0;
  uint8_t* totals = nullptr;  // pinned: every position's region total
  uint64_t totals_cap = 0;
  std::vector<uint32_t> totals_h;
};

// Pinned staging that grows by at least half; the device is idle before the old
// block is freed, since an earlier restart's copy may have used it.
static void pinned_reserve(uint8_t*& p, uint64_t& cap, uint64_t want) {
  if (want <= cap) return;
  const uint64_t c = std::max<uint64_t>({want, cap + cap / 2, 4096});
  if (p) {
    MTC_HIP(hipDeviceSynchronize());
    (void)hipHostFree(p);
  }
  p = nullptr;
  cap = 0;
  MTC_HIP(hipHostMalloc((void**)&p, c));
  cap = c;
}

static std::mutex& scratch_mutex() {
  static std::mutex m;
  return m;
}

// Scratch of one device, kept until the process ends.
static RestartScratch& scratch_for(int dev) {
  static std::map<int, RestartScratch*> all;
  RestartScratch*& r = all[dev];
  if (!r) r = new RestartScratch();
  return *r;
}

// The calling thread's last restart: upload of a host chain, and the whole restart
// (restart_timers of src/tree_approach.cpp:1792-1793).
static thread_local float g_restart_ms[2] = {0, 0};

void last_restart_times(float* upload_ms, float* total_ms) {
  if (upload_ms) *upload_ms = g_restart_ms[0];
  if (total_ms) *total_ms = g_restart_ms[1];
}

static void check_err(uint32_t* err_d, hipStream_t st) {
  uint32_t e = 0;
  MTC_HIP(hipMemcpyAsync(&e, err_d, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  MTC_HIP(hipStreamSynchronize(st));
  if (e & ERR_TABLE_FULL) throw Status{MTC_ERR_OUT_OF_MEMORY, "restart root set overflow"};
  if (e) throw Status{MTC_ERR_CORRUPT_DIFF, "diff chain is inconsistent (err bits " + std::to_string(e) + ")"};
}

void restart_impl(const void* const* diffs, const uint64_t* sizes, uint32_t count, uint32_t k, bool on_device,
                  uint8_t* out, uint64_t out_len, hipStream_t st, bool basic) {
  if (!diffs || !sizes || k >= count || !diffs[k]) throw Status{MTC_ERR_INVALID_ARG, "bad diff chain arguments"};
  // Scratch buffers are kept per device between restarts (grow-only); one restart
  // at a time uses them.
  int dev = 0;
  MTC_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> guard(scratch_mutex());
  RestartScratch& rs = scratch_for(dev);
  g_restart_ms[0] = g_restart_ms[1] = 0;
  if (!rs.ev[0])
    for (auto& e : rs.ev) MTC_HIP(hipEventCreate(&e));
  MTC_HIP(hipEventRecord(rs.ev[0], st));
  auto present = [&](uint32_t idx) {
    if (!diffs[idx] || sizes[idx] < sizeof(mtc_header)) throw Status{MTC_ERR_CORRUPT_DIFF, "missing or short diff"};
  };
  // Headers of device chains come back through pinned staging, one
  // synchronisation per batch of copies.
  auto read_headers = [&](uint32_t first, uint32_t num, mtc_header* out) {
    for (uint32_t i = 0; i < num; i++) present(first + i);
    if (!on_device) {
      for (uint32_t i = 0; i < num; i++) std::memcpy(out + i, diffs[first + i], sizeof(mtc_header));
      return;
    }
    pinned_reserve(rs.hstage, rs.hstage_cap, (uint64_t)num * sizeof(mtc_header));
    for (uint32_t i = 0; i < num; i++)
      MTC_HIP(hipMemcpyAsync(rs.hstage + (uint64_t)i * sizeof(mtc_header), diffs[first + i], sizeof(mtc_header),
                             hipMemcpyDeviceToHost, st));
    MTC_HIP(hipStreamSynchronize(st));
    std::memcpy(out, rs.hstage, (uint64_t)num * sizeof(mtc_header));
  };
  mtc_header hk;
  read_headers(k, 1, &hk);
  if (hk.chkpt_id != k || hk.ref_id > k || hk.chunk_size == 0 || hk.datalen == 0)
    throw Status{MTC_ERR_CORRUPT_DIFF, "bad header of the requested diff"};
  if (out_len < hk.datalen) throw Status{MTC_ERR_BUFFER_TOO_SMALL, "output shorter than datalen"};
  const uint64_t C = hk.chunk_size;
  const uint64_t n64 = (hk.datalen + C - 1) / C;
  if (n64 > (1ull << 31)) throw Status{MTC_ERR_CORRUPT_DIFF, "too many chunks"};
  const uint32_t n = (uint32_t)n64, ref = hk.ref_id, L = k - ref + 1;
  const Geo g = make_geo(n);
  const uint32_t idb = basic ? n - 1 : 0;

  // Chain layout, with host diffs uploaded once.
  std::vector<ChainDiff> ch(L);
  {
    std::vector<mtc_header> hs(L);
    read_headers(ref, L, hs.data());
    for (uint32_t p = 0; p < L; p++) ch[p].h = hs[p];
  }
  uint64_t up_total = 0;
  if (!on_device)
    for (uint32_t p = 0; p < L; p++) up_total += (sizes[ref + p] + 255) & ~255ull;
  // host chains: an exact-size upload, freed when the restart returns (a chain
  // can be as large as the buffer, so it is not kept)
  DevBuf<uint8_t> upload;
  if (!on_device) upload.ensure(up_total);
  uint64_t up_off = 0, sumF = 0, pt_total = 0;
  for (uint32_t p = 0; p < L; p++) {
    const uint32_t idx = ref + p;
    ChainDiff& d = ch[p];
    if (d.h.chkpt_id != idx || d.h.ref_id != ref || d.h.datalen != hk.datalen || d.h.chunk_size != hk.chunk_size)
      throw Status{MTC_ERR_CORRUPT_DIFF, "diff " + std::to_string(idx) + " does not belong to the chain"};
    if (basic && (d.h.num_prior_chkpts || d.h.num_shift_dupl))
      throw Status{MTC_ERR_CORRUPT_DIFF, "diff " + std::to_string(idx) + " is not a basic diff"};
    d.size = sizes[idx];
    d.off_first = 32;
    d.off_table = d.off_first + 4ull * d.h.num_first_ocur;
    d.off_shift = d.off_table + 8ull * d.h.num_prior_chkpts;
    d.off_data = d.off_shift + 8ull * d.h.num_shift_dupl;
    if (d.off_data > d.size) throw Status{MTC_ERR_CORRUPT_DIFF, "diff shorter than its metadata"};
    d.data_chunks = (d.size - d.off_data) / C;
    if (on_device) {
      d.base = static_cast<const uint8_t*>(diffs[idx]);
    } else {
      MTC_HIP(hipMemcpyAsync(upload.p + up_off, diffs[idx], d.size, hipMemcpyHostToDevice, st));
      d.base = upload.p + up_off;
      up_off += (d.size + 255) & ~255ull;
    }
    d.region_base = sumF + p;
    sumF += d.h.num_first_ocur;
    pt_total += 8ull * d.h.num_prior_chkpts;
  }
  MTC_HIP(hipEventRecord(rs.ev[1], st));
  // The prior-checkpoint tables (in bounds: off_data <= size was checked above).
  std::vector<std::vector<uint32_t>> ptables(L);
  if (on_device && pt_total) pinned_reserve(rs.hstage, rs.hstage_cap, pt_total);
  uint64_t pt_off = 0;
  for (uint32_t p = 0; p < L; p++) {
    const ChainDiff& d = ch[p];
    const uint64_t bytes = 8ull * d.h.num_prior_chkpts;
    ptables[p].resize(2ull * d.h.num_prior_chkpts);
    if (!bytes) continue;
    if (on_device)
      MTC_HIP(hipMemcpyAsync(rs.hstage + pt_off, d.base + d.off_table, bytes, hipMemcpyDeviceToHost, st));
    else
      std::memcpy(ptables[p].data(), static_cast<const uint8_t*>(diffs[ref + p]) + d.off_table, bytes);
    pt_off += bytes;
  }
  if (on_device && pt_total) {
    MTC_HIP(hipStreamSynchronize(st));
    pt_off = 0;
    for (uint32_t p = 0; p < L; p++) {
      const uint64_t bytes = 8ull * ch[p].h.num_prior_chkpts;
      if (bytes) std::memcpy(ptables[p].data(), rs.hstage + pt_off, bytes);
      pt_off += bytes;
    }
  }

  Workspace& ws = rs.ws;
  rs.err.ensure(1);
  k_r_fill32<<<1, kRThreads, 0, st>>>(rs.err.p, 1, 0);
  rs.region_off.reserve(sumF + L);
  uint64_t set_cap = 16;
  while (set_cap < 2 * sumF) set_cap <<= 1;
  rs.set_keys.ensure(set_cap);
  rs.set_vals.ensure(set_cap);
  k_r_fill64<<<rgrid(set_cap, 65536), kRThreads, 0, st>>>(rs.set_keys.p, set_cap, kNoCover);
  rs.cover.ensure(n);
  k_r_fill64<<<rgrid(n, 65536), kRThreads, 0, st>>>(rs.cover.p, n, kNoCover);
  MTC_HIP(hipGetLastError());
  uint32_t* err_p = rs.err.p;
  uint32_t* region_p = rs.region_off.p;
  uint64_t* cover_p = rs.cover.p;

  // Per-position tables for the kernels: F-root address, region base, data address,
  // data chunks; then every position's running count of shifted entries by source
  // checkpoint (ascending id, :872-883).
  std::vector<uint64_t> meta_h(4ull * L);
  for (uint32_t p = 0; p < L; p++) {
    meta_h[p] = reinterpret_cast<uint64_t>(ch[p].base + ch[p].off_first);
    meta_h[L + p] = ch[p].region_base;
    meta_h[2 * L + p] = reinterpret_cast<uint64_t>(ch[p].base + ch[p].off_data);
    meta_h[3 * L + p] = ch[p].data_chunks;
  }
  std::vector<uint32_t> cum_h;
  std::vector<uint64_t> cum_off(L, 0);
  for (uint32_t p = 0; p < L; p++) {
    const uint32_t idx = ref + p, S = ch[p].h.num_shift_dupl;
    cum_off[p] = cum_h.size();
    std::vector<uint32_t> cnt(idx - ref + 1, 0);
    for (uint32_t t = 0; t < ch[p].h.num_prior_chkpts; t++) {
      const uint32_t c = ptables[p][2 * t], m = ptables[p][2 * t + 1];
      if (c < ref || c > idx) throw Status{MTC_ERR_CORRUPT_DIFF, "shifted entries name a checkpoint off the chain"};
      cnt[c - ref] += m;
    }
    for (size_t t = 1; t < cnt.size(); t++) cnt[t] += cnt[t - 1];
    if (cnt.back() != S) throw Status{MTC_ERR_CORRUPT_DIFF, "prior-checkpoint table does not sum to S"};
    cum_h.insert(cum_h.end(), cnt.begin(), cnt.end());
  }
  // meta and cum share one upload from pinned staging
  const uint64_t meta_bytes = 8 * meta_h.size(), cum_bytes = 4 * cum_h.size();
  rs.meta.reserve(meta_h.size() + (cum_h.size() + 1) / 2);
  pinned_reserve(rs.stage, rs.stage_cap, meta_bytes + cum_bytes);
  MTC_HIP(hipStreamSynchronize(st));  // the staging may still feed the previous restart's copy
  std::memcpy(rs.stage, meta_h.data(), meta_bytes);
  std::memcpy(rs.stage + meta_bytes, cum_h.data(), cum_bytes);
  MTC_HIP(hipMemcpyAsync(rs.meta.p, rs.stage, meta_bytes + cum_bytes, hipMemcpyHostToDevice, st));
  const uint64_t* meta_p = rs.meta.p;
  const uint32_t* cum_p = reinterpret_cast<const uint32_t*>(rs.meta.p + meta_h.size());

  // Region offsets and the (position, root) set.
  for (uint32_t p = 0; p < L; p++) {
    const ChainDiff& d = ch[p];
    const uint32_t F = d.h.num_first_ocur;
    uint32_t* ro = region_p + d.region_base;
    if (F) {
      k_r_roots<<<rgrid(F), kRThreads, 0, st>>>(d.base + d.off_first, F, g, ro, p, rs.set_keys.p, rs.set_vals.p,
                                                set_cap - 1, err_p, idb);
      MTC_HIP(hipGetLastError());
    }
    exclusive_scan_u32(ro, ro, F, ro + F, ws, st);
  }
  rs.totals_h.resize(L);
  // Each position's total sits at region_base + F.  A short region array comes
  // back in one copy; a long one total by total.
  const uint64_t region_words = sumF + L;
  const bool whole = region_words <= (1u << 20);
  pinned_reserve(rs.totals, rs.totals_cap, 4ull * (whole ? region_words : L));
  uint32_t* totals = reinterpret_cast<uint32_t*>(rs.totals);
  if (whole) {
    MTC_HIP(hipMemcpyAsync(totals, region_p, 4ull * region_words, hipMemcpyDeviceToHost, st));
  } else {
    for (uint32_t p = 0; p < L; p++)
      MTC_HIP(hipMemcpyAsync(totals + p, region_p + ch[p].region_base + ch[p].h.num_first_ocur, sizeof(uint32_t),
                             hipMemcpyDeviceToHost, st));
  }
  check_err(err_p, st);
  for (uint32_t p = 0; p < L; p++) {
    rs.totals_h[p] = whole ? totals[ch[p].region_base + ch[p].h.num_first_ocur] : totals[p];
    if (rs.totals_h[p] > ch[p].data_chunks) throw Status{MTC_ERR_CORRUPT_DIFF, "data section shorter than its regions"};
  }

  // Cover chunks, oldest diff first so newer diffs overwrite.
  for (uint32_t p = 0; p < L; p++) {
    const ChainDiff& d = ch[p];
    const uint32_t F = d.h.num_first_ocur, S = d.h.num_shift_dupl;
    if (F) {
      k_r_cover_first<<<rgrid(rs.totals_h[p], 65536), kRThreads, 0, st>>>(d.base + d.off_first, F,
                                                                         region_p + d.region_base, rs.totals_h[p], g, p,
                                                                         cover_p, idb);
      MTC_HIP(hipGetLastError());
    }
    if (S) {
      RShiftArgs sa{d.base + d.off_shift, S, cum_p + cum_off[p], p + 1, ref, p, g, rs.set_keys.p, rs.set_vals.p,
                   set_cap - 1, meta_p, meta_p + L, region_p, cover_p, err_p};
      k_r_cover_shift<<<rgrid(S), kRThreads, 0, st>>>(sa);
      MTC_HIP(hipGetLastError());
    }
  }
  const int vec = (C % 16 == 0) && ((uintptr_t)out % 16 == 0);
  k_r_copy<<<rgrid(n, 8192), kRThreads, 0, st>>>(cover_p, n, (uint32_t)C, hk.datalen, meta_p + 2 * L, meta_p + 3 * L,
                                                 out, vec, err_p);
  MTC_HIP(hipGetLastError());
  MTC_HIP(hipEventRecord(rs.ev[2], st));
  check_err(err_p, st);
  MTC_HIP(hipEventElapsedTime(&g_restart_ms[0], rs.ev[0], rs.ev[1]));
  MTC_HIP(hipEventElapsedTime(&g_restart_ms[1], rs.ev[0], rs.ev[2]));
  if (on_device) g_restart_ms[0] = 0;  // nothing was uploaded
}

// ---------------------------------------------------------------------------
// Restart without allocation (mtc_restart_noalloc).  Every scratch buffer of the
// restart above -- the cover array, the (position, root) set, region offsets, the
// per-position tables and scan partials -- is carved from one static device array
// per GPU, and the host side uses no pinned staging: headers come back through one
// gather kernel and one small copy, and the running counts of shifted entries by
// source checkpoint are built on the device.  Chains are device-resident tree
// diffs of at most kNoAllocMaxChain checkpoints.
// ---------------------------------------------------------------------------
constexpr uint64_t kNoAllocScratch = 64ull << 20;  // bytes of `scratch` on each GPU
constexpr uint32_t kNoAllocMaxChain = 256;         // diffs from ref_id to the restarted checkpoint
constexpr int kNoAllocMaxDevices = 64;

static __device__ __align__(256) uint8_t scratch[kNoAllocScratch];

// Per GPU: the device address of `scratch` (taken once, with that GPU current) and
// the lock that keeps restarts on one GPU from sharing the array at the same time.
struct NoAllocDevice {
  std::mutex mu;
  uint8_t* ptr = nullptr;
};
static NoAllocDevice g_noalloc[kNoAllocMaxDevices];

// The 32-byte headers of the chain's diffs, gathered next to each other.
__global__ void k_r_gather_headers(const uint64_t* ptrs, uint32_t L, uint32_t* out) {
  const uint32_t i = blockIdx.x * kRThreads + threadIdx.x;
  if (i >= 8 * L) return;
  out[i] = ld32(reinterpret_cast<const uint8_t*>(ptrs[i / 8]) + 4 * (i % 8));
}

// One block per chain position p: the inclusive running count, by checkpoint id
// ref..ref+p, of the shifted entries of diff p (its prior-checkpoint table,
// src/tree_approach.cpp:872-883), checked against the diff's S.
__global__ __launch_bounds__(kRThreads) void k_r_cum(const uint64_t* table_ptr, const uint64_t* ps,
                                                     const uint64_t* cum_off, uint32_t ref, uint32_t* cum,
                                                     uint32_t* err) {
  __shared__ uint32_t cnt[kNoAllocMaxChain];
  const uint32_t p = blockIdx.x, t = threadIdx.x;
  const uint32_t P = (uint32_t)(ps[p] >> 32), S = (uint32_t)ps[p];
  const uint8_t* tab = reinterpret_cast<const uint8_t*>(table_ptr[p]);
  cnt[t] = 0;
  __syncthreads();
  for (uint32_t i = t; i < P; i += kRThreads) {
    const uint32_t c = ld32(tab + 8ull * i), m = ld32(tab + 8ull * i + 4);
    if (c < ref || c > ref + p) atomicOr(err, ERR_CORRUPT);
    else atomicAdd(&cnt[c - ref], m);
  }
  __syncthreads();
  for (uint32_t o = 1; o < kNoAllocMaxChain; o <<= 1) {  // inclusive scan
    const uint32_t v = (t >= o) ? cnt[t - o] : 0u;
    __syncthreads();
    cnt[t] += v;
    __syncthreads();
  }
  if (t <= p) cum[cum_off[p] + t] = cnt[t];
  if (t == p && cnt[t] != S) atomicOr(err, ERR_CORRUPT);
}

// k_r_cover_first with the region total read on the device (region_off[F]) and
// checked against the data section's capacity.
__global__ void k_r_cover_first_dev(const uint8_t* first, uint32_t F, const uint32_t* region_off,
                                    uint64_t data_chunks, Geo g, uint32_t pos, uint64_t* cover, uint32_t* err) {
  const uint64_t total = region_off[F];
  if (total > data_chunks) {
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(err, ERR_CORRUPT);
    return;
  }
  for (uint64_t j = (uint64_t)blockIdx.x * kRThreads + threadIdx.x; j < total; j += (uint64_t)gridDim.x * kRThreads) {
    uint32_t lo = 0, hi = F;
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (region_off[mid] <= j) lo = mid;
      else hi = mid;
    }
    uint32_t st[2], ct[2];
    const int r = leaf_runs(g, ld32(first + 4ull * lo), st, ct);
    const uint32_t chunk = run_at(r, st, ct, (uint32_t)(j - region_off[lo]));
    cover[chunk] = ((uint64_t)pos << kSlotBits) | j;
  }
}

// Bump carving of the scratch array (256-byte aligned pieces).
struct Carve {
  uint64_t off = 0;
  uint64_t take(uint64_t bytes) {
    const uint64_t o = off;
    off = (off + bytes + 255) & ~255ull;
    return o;
  }
};

void restart_noalloc_impl(const void* const* diffs, const uint64_t* sizes, uint32_t count, uint32_t k, uint8_t* out,
                          uint64_t out_len, hipStream_t st) {
  if (!diffs || !sizes || k >= count || !diffs[k]) throw Status{MTC_ERR_INVALID_ARG, "bad diff chain arguments"};
  int dev = 0;
  MTC_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNoAllocMaxDevices) throw Status{MTC_ERR_INVALID_ARG, "device ordinal above 63"};
  NoAllocDevice& nd = g_noalloc[dev];
  std::lock_guard<std::mutex> guard(nd.mu);
  if (!nd.ptr) {  // once per GPU, with it current: the device copy of `scratch`
    void* p = nullptr;
    MTC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw Status{MTC_ERR_HIP, "no device address for the restart scratch"};
    nd.ptr = static_cast<uint8_t*>(p);
  }
  uint8_t* const base = nd.ptr;
  g_restart_ms[0] = g_restart_ms[1] = 0;
  if (sizes[k] < sizeof(mtc_header)) throw Status{MTC_ERR_CORRUPT_DIFF, "missing or short diff"};
  mtc_header hk;
  MTC_HIP(hipMemcpyAsync(&hk, diffs[k], sizeof(hk), hipMemcpyDeviceToHost, st));
  MTC_HIP(hipStreamSynchronize(st));
  if (hk.chkpt_id != k || hk.ref_id > k || hk.chunk_size == 0 || hk.datalen == 0)
    throw Status{MTC_ERR_CORRUPT_DIFF, "bad header of the requested diff"};
  if (out_len < hk.datalen) throw Status{MTC_ERR_BUFFER_TOO_SMALL, "output shorter than datalen"};
  const uint64_t C = hk.chunk_size;
  const uint64_t n64 = (hk.datalen + C - 1) / C;
  if (n64 > (1ull << 31)) throw Status{MTC_ERR_CORRUPT_DIFF, "too many chunks"};
  const uint32_t n = (uint32_t)n64, ref = hk.ref_id, L = k - ref + 1;
  if (L > kNoAllocMaxChain)
    throw Status{MTC_ERR_INVALID_ARG, "the no-allocation restart takes chains of at most 256 diffs"};
  const Geo g = make_geo(n);

  // Headers of the chain: one gather kernel, one copy.
  Carve cv;
  uint32_t* err_p = reinterpret_cast<uint32_t*>(base + cv.take(4));
  uint64_t* ptrs_d = reinterpret_cast<uint64_t*>(base + cv.take(8ull * L));
  uint32_t* hdr_d = reinterpret_cast<uint32_t*>(base + cv.take(32ull * L));
  uint64_t ptrs_h[kNoAllocMaxChain];
  for (uint32_t p = 0; p < L; p++) {
    const uint32_t idx = ref + p;
    if (!diffs[idx] || sizes[idx] < sizeof(mtc_header)) throw Status{MTC_ERR_CORRUPT_DIFF, "missing or short diff"};
    ptrs_h[p] = reinterpret_cast<uint64_t>(diffs[idx]);
  }
  MTC_HIP(hipMemcpyAsync(ptrs_d, ptrs_h, 8ull * L, hipMemcpyHostToDevice, st));
  k_r_gather_headers<<<rgrid(8ull * L), kRThreads, 0, st>>>(ptrs_d, L, hdr_d);
  MTC_HIP(hipGetLastError());
  mtc_header hs[kNoAllocMaxChain];
  MTC_HIP(hipMemcpyAsync(hs, hdr_d, 32ull * L, hipMemcpyDeviceToHost, st));
  MTC_HIP(hipStreamSynchronize(st));

  // Chain layout (the same checks as the allocating restart).
  uint64_t off_first[kNoAllocMaxChain], off_shift[kNoAllocMaxChain], off_table[kNoAllocMaxChain];
  uint64_t meta_h[7 * kNoAllocMaxChain];  // F-root, region base, data, data chunks, table, (P, S), cum offset
  uint64_t sumF = 0, cum_words = 0;
  uint32_t maxF = 0;
  for (uint32_t p = 0; p < L; p++) {
    const uint32_t idx = ref + p;
    const mtc_header& h = hs[p];
    if (h.chkpt_id != idx || h.ref_id != ref || h.datalen != hk.datalen || h.chunk_size != hk.chunk_size)
      throw Status{MTC_ERR_CORRUPT_DIFF, "diff " + std::to_string(idx) + " does not belong to the chain"};
    off_first[p] = 32;
    off_table[p] = off_first[p] + 4ull * h.num_first_ocur;
    off_shift[p] = off_table[p] + 8ull * h.num_prior_chkpts;
    const uint64_t off_data = off_shift[p] + 8ull * h.num_shift_dupl;
    if (off_data > sizes[idx]) throw Status{MTC_ERR_CORRUPT_DIFF, "diff shorter than its metadata"};
    if (h.num_prior_chkpts > p + 1) throw Status{MTC_ERR_CORRUPT_DIFF, "prior-checkpoint table longer than the chain"};
    const uint8_t* b = static_cast<const uint8_t*>(diffs[idx]);
    meta_h[p] = reinterpret_cast<uint64_t>(b + off_first[p]);
    meta_h[L + p] = sumF + p;
    meta_h[2 * L + p] = reinterpret_cast<uint64_t>(b + off_data);
    meta_h[3 * L + p] = (sizes[idx] - off_data) / C;
    meta_h[4 * L + p] = reinterpret_cast<uint64_t>(b + off_table[p]);
    meta_h[5 * L + p] = ((uint64_t)h.num_prior_chkpts << 32) | h.num_shift_dupl;
    meta_h[6 * L + p] = cum_words;
    cum_words += p + 1;
    sumF += h.num_first_ocur;
    maxF = std::max(maxF, h.num_first_ocur);
  }
  uint64_t set_cap = 16;
  while (set_cap < 2 * sumF) set_cap <<= 1;
  uint64_t* cover_p = reinterpret_cast<uint64_t*>(base + cv.take(8ull * n));
  uint64_t* keys_p = reinterpret_cast<uint64_t*>(base + cv.take(8 * set_cap));
  uint32_t* vals_p = reinterpret_cast<uint32_t*>(base + cv.take(4 * set_cap));
  const uint64_t off_region = cv.take(4 * (sumF + L));
  const uint64_t off_meta = cv.take(8ull * 7 * L);
  const uint64_t off_cum = cv.take(4 * cum_words);
  uint64_t off_sums[3];
  for (int d = 0; d < 3; d++) off_sums[d] = cv.take(4 * scan_sums_words(maxF, d));
  if (cv.off > kNoAllocScratch)
    throw Status{MTC_ERR_INVALID_ARG, "this restart needs " + std::to_string(cv.off) +
                                          " bytes of scratch; the no-allocation restart has 64 MiB per GPU"};
  uint32_t* region_p = reinterpret_cast<uint32_t*>(base + off_region);
  uint64_t* meta_p = reinterpret_cast<uint64_t*>(base + off_meta);
  uint32_t* cum_p = reinterpret_cast<uint32_t*>(base + off_cum);
  uint32_t* const sums[3] = {reinterpret_cast<uint32_t*>(base + off_sums[0]),
                             reinterpret_cast<uint32_t*>(base + off_sums[1]),
                             reinterpret_cast<uint32_t*>(base + off_sums[2])};
  MTC_HIP(hipMemcpyAsync(meta_p, meta_h, 8ull * 7 * L, hipMemcpyHostToDevice, st));
  // The array keeps the previous call's contents: everything read before it is
  // written starts from a fill.
  k_r_fill32<<<1, kRThreads, 0, st>>>(err_p, 1, 0);
  k_r_fill64<<<rgrid(set_cap, 65536), kRThreads, 0, st>>>(keys_p, set_cap, kNoCover);
  k_r_fill64<<<rgrid(n, 65536), kRThreads, 0, st>>>(cover_p, n, kNoCover);
  k_r_cum<<<L, kRThreads, 0, st>>>(meta_p + 4 * L, meta_p + 5 * L, meta_p + 6 * L, ref, cum_p, err_p);
  MTC_HIP(hipGetLastError());
  for (uint32_t p = 0; p < L; p++) {
    const uint32_t F = hs[p].num_first_ocur;
    const uint8_t* b = static_cast<const uint8_t*>(diffs[ref + p]);
    uint32_t* ro = region_p + meta_h[L + p];
    if (F) {
      k_r_roots<<<rgrid(F), kRThreads, 0, st>>>(b + off_first[p], F, g, ro, p, keys_p, vals_p, set_cap - 1, err_p, 0);
      MTC_HIP(hipGetLastError());
    }
    exclusive_scan_u32_fixed(ro, ro, F, ro + F, sums, st);
  }
  check_err(err_p, st);  // every root lies inside the tree before any cover is written
  for (uint32_t p = 0; p < L; p++) {
    const uint32_t F = hs[p].num_first_ocur, S = hs[p].num_shift_dupl;
    const uint8_t* b = static_cast<const uint8_t*>(diffs[ref + p]);
    if (F) {
      k_r_cover_first_dev<<<rgrid(std::max<uint64_t>(meta_h[3 * L + p], 1), 65536), kRThreads, 0, st>>>(
          b + off_first[p], F, region_p + meta_h[L + p], meta_h[3 * L + p], g, p, cover_p, err_p);
      MTC_HIP(hipGetLastError());
    }
    if (S) {
      RShiftArgs sa{b + off_shift[p], S, cum_p + meta_h[6 * L + p], p + 1, ref, p, g, keys_p, vals_p, set_cap - 1,
                   meta_p, meta_p + L, region_p, cover_p, err_p};
      k_r_cover_shift<<<rgrid(S), kRThreads, 0, st>>>(sa);
      MTC_HIP(hipGetLastError());
    }
  }
  const int vec = (C % 16 == 0) && ((uintptr_t)out % 16 == 0);
  k_r_copy<<<rgrid(n, 8192), kRThreads, 0, st>>>(cover_p, n, (uint32_t)C, hk.datalen, meta_p + 2 * L, meta_p + 3 * L,
                                                 out, vec, err_p);
  MTC_HIP(hipGetLastError());
  check_err(err_p, st);
}

}  // namespace mtc
