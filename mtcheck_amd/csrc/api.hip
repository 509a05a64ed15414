// api.hip -- the C ABI of include/mtcheck.h over the HIP implementation.
#include <cstring>
#include <new>

#include "internal.hpp"

namespace mtc {
void checkpoint_impl(mtc_dedup* h, const uint8_t* data, uint64_t len, bool make_baseline, hipStream_t st);
void restart_impl(const void* const* diffs, const uint64_t* sizes, uint32_t count, uint32_t k, bool on_device,
                  uint8_t* out, uint64_t out_len, hipStream_t st, bool basic);
void last_restart_times(float* upload_ms, float* total_ms);
void restart_noalloc_impl(const void* const* diffs, const uint64_t* sizes, uint32_t count, uint32_t k, uint8_t* out,
                          uint64_t out_len, hipStream_t st);
void hash_chunks(const uint8_t* data, uint64_t len, uint32_t C, int hk, Digest* out, hipStream_t st);
void checkpoint_batch_impl(mtc_dedup* const* hs, const uint8_t* const* datas, const uint64_t* lens,
                           const int* baselines, uint32_t count, StepCtx& ctx, hipStream_t st, bool use_graph);
}  // namespace mtc

// A batch context: its own stream (default priority) and step state.
struct mtc_batch {
  int device = 0;
  int use_graph = 1;
  hipStream_t stream = nullptr;
  hipEvent_t ready = nullptr;
  mtc::StepCtx step;
  std::string last_error;
};

using mtc::HipError;
using mtc::Status;

static thread_local std::string g_last_error;

template <class F>
static int guarded(mtc_dedup* h, F&& body) {
  std::string msg;
  int code = MTC_OK;
  try {
    body();
  } catch (const Status& s) {
    code = s.code;
    msg = s.msg;
  } catch (const HipError& e) {
    code = (e.err == hipErrorOutOfMemory) ? MTC_ERR_OUT_OF_MEMORY : MTC_ERR_HIP;
    msg = std::string(e.what) + ": " + hipGetErrorString(e.err);
    (void)hipGetLastError();
  } catch (const std::bad_alloc&) {
    code = MTC_ERR_OUT_OF_MEMORY;
    msg = "host allocation failed";
  } catch (...) {
    code = MTC_ERR_HIP;
    msg = "unexpected exception";
  }
  if (code != MTC_OK) {
    g_last_error = msg;
    if (h) h->last_error = msg;
  }
  return code;
}

static void use_device(int dev) {
  int cnt = 0;
  MTC_HIP(hipGetDeviceCount(&cnt));
  if (dev < 0 || dev >= cnt) throw Status{MTC_ERR_INVALID_ARG, "no such HIP device"};
  MTC_HIP(hipSetDevice(dev));
}

extern "C" {

int mtc_create(mtc_dedup** out, uint32_t chunk_size, int approach, int hash_kind, int device) {
  if (!out) return MTC_ERR_INVALID_ARG;
  *out = nullptr;
  mtc_dedup* h = nullptr;
  int rc = guarded(nullptr, [&] {
    if (chunk_size == 0) throw Status{MTC_ERR_INVALID_ARG, "chunk size must be positive"};
    if (hash_kind != MTC_HASH_MURMUR3 && hash_kind != MTC_HASH_MD5) throw Status{MTC_ERR_INVALID_ARG, "unknown hash"};
    switch (approach) {
      case MTC_APPROACH_UNKNOWN:  // the reference falls through to the tree path
      case MTC_APPROACH_TREE:
      case MTC_APPROACH_TREE_LOW_OFFSET_REF:
      case MTC_APPROACH_TREE_LOW_OFFSET:
      case MTC_APPROACH_TREE_LOW_ROOT_REF:
      case MTC_APPROACH_TREE_LOW_ROOT:
      case MTC_APPROACH_BASIC:
        break;
      case MTC_APPROACH_FULL:
      case MTC_APPROACH_LIST:
        throw Status{MTC_ERR_UNSUPPORTED_APPROACH, "the tree approach is implemented"};
      default:
        throw Status{MTC_ERR_INVALID_ARG, "unknown approach"};
    }
    use_device(device);
    h = new mtc_dedup();
    h->device = device;
    h->chunk_size = chunk_size;
    h->hash_kind = hash_kind;
    h->approach = approach;
    MTC_HIP(hipMalloc((void**)&h->ctl_d, sizeof(mtc::Ctl)));
    MTC_HIP(hipHostMalloc((void**)&h->ctl_h, sizeof(mtc::Ctl), hipHostMallocDefault));
  });
  if (rc != MTC_OK) {
    if (h) mtc_destroy(h);
    return rc;
  }
  *out = h;
  return MTC_OK;
}

int mtc_destroy(mtc_dedup* h) {
  if (!h) return MTC_OK;
  (void)hipSetDevice(h->device);
  if (h->copy_st) {
    (void)hipStreamSynchronize(h->copy_st);
    (void)hipStreamDestroy(h->copy_st);
    for (hipEvent_t e : h->copy_done)
      if (e) (void)hipEventDestroy(e);
  }
  if (h->ctl_d) (void)hipFree(h->ctl_d);
  if (h->ctl_h) (void)hipHostFree(h->ctl_h);
  delete h;  // device buffers are released by their destructors
  return MTC_OK;
}

int mtc_checkpoint(mtc_dedup* h, const void* dev_data, uint64_t len, int make_baseline, int advance_id, void* stream,
                   mtc_header* header_out, uint64_t* diff_size_out) {
  if (!h) return MTC_ERR_INVALID_ARG;
  return guarded(h, [&] {
    if (!dev_data || len == 0) throw Status{MTC_ERR_INVALID_ARG, "empty input"};
    MTC_HIP(hipSetDevice(h->device));
    mtc::checkpoint_impl(h, static_cast<const uint8_t*>(dev_data), len, make_baseline != 0,
                         static_cast<hipStream_t>(stream));
    if (header_out) *header_out = h->last_header;
    if (diff_size_out) *diff_size_out = h->last_diff_size;
    if (advance_id) h->current_id += 1;
  });
}

int mtc_diff_device(mtc_dedup* h, const void** dev_diff, uint64_t* size) {
  if (!h || !dev_diff || !size) return MTC_ERR_INVALID_ARG;
  if (!h->has_state || !h->last_diff_size) return MTC_ERR_STATE;
  *dev_diff = h->diff_ptr();
  *size = h->last_diff_size;
  return MTC_OK;
}

int mtc_copy_diff(mtc_dedup* h, void* host_dst, uint64_t capacity, void* stream) {
  if (!h || !host_dst) return MTC_ERR_INVALID_ARG;
  return guarded(h, [&] {
    if (!h->last_diff_size) throw Status{MTC_ERR_STATE, "no checkpoint taken"};
    if (capacity < h->last_diff_size) throw Status{MTC_ERR_BUFFER_TOO_SMALL, "host buffer smaller than the diff"};
    MTC_HIP(hipSetDevice(h->device));
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipEvent_t a, b;
    MTC_HIP(hipEventCreate(&a));
    MTC_HIP(hipEventCreate(&b));
    MTC_HIP(hipEventRecord(a, st));
    // host or device destination (unified addressing decides the direction)
    MTC_HIP(hipMemcpyAsync(host_dst, h->diff_ptr(), h->last_diff_size, hipMemcpyDefault, st));
    MTC_HIP(hipEventRecord(b, st));
    MTC_HIP(hipEventSynchronize(b));
    (void)hipEventElapsedTime(&h->last_ms[3], a, b);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
  });
}

int mtc_copy_diff_async(mtc_dedup* h, void* host_dst, uint64_t capacity) {
  if (!h || !host_dst) return MTC_ERR_INVALID_ARG;
  return guarded(h, [&] {
    if (!h->last_diff_size) throw Status{MTC_ERR_STATE, "no checkpoint taken"};
    if (capacity < h->last_diff_size) throw Status{MTC_ERR_BUFFER_TOO_SMALL, "host buffer smaller than the diff"};
    MTC_HIP(hipSetDevice(h->device));
    if (!h->copy_st) {
      MTC_HIP(hipStreamCreateWithFlags(&h->copy_st, hipStreamNonBlocking));
      for (int i = 0; i < 2; i++) MTC_HIP(hipEventCreateWithFlags(&h->copy_done[i], hipEventDisableTiming));
    }
    h->async_copies = true;  // from the next checkpoint on, checkpoints alternate diff arenas
    // The second arena is allocated here, before the copy is queued, while the
    // device is idle (the checkpoint has returned), rather than inside the next
    // checkpoint.
    h->diff_alt.ensure(h->diff.count);
    // The diff is complete: mtc_checkpoint returns after its last kernel.
    MTC_HIP(hipMemcpyAsync(host_dst, h->diff_ptr(), h->last_diff_size, hipMemcpyDefault, h->copy_st));
    MTC_HIP(hipEventRecord(h->copy_done[h->diff_cur], h->copy_st));
    h->copy_pending[h->diff_cur] = true;
  });
}

int mtc_copy_wait(mtc_dedup* h) {
  if (!h) return MTC_ERR_INVALID_ARG;
  return guarded(h, [&] {
    if (h->copy_st) MTC_HIP(hipStreamSynchronize(h->copy_st));
  });
}

int mtc_get_stats(const mtc_dedup* h, mtc_stats* out) {
  if (!h || !out) return MTC_ERR_INVALID_ARG;
  std::memset(out, 0, sizeof(*out));
  out->current_id = h->current_id;
  out->baseline_id = h->baseline_id;
  out->num_first_ocur = h->last_header.num_first_ocur;
  out->num_shift_dupl = h->last_header.num_shift_dupl;
  out->num_prior_chkpts = h->last_header.num_prior_chkpts;
  out->num_chunks = h->n;
  out->data_len = h->data_len;
  out->diff_size = h->last_diff_size;
  out->data_chunks = h->ctl_h ? h->ctl_h->data_chunks : 0;
  out->table_entries = h->table_size;
  out->table_capacity = h->table_cap;
  out->ms_setup = h->last_ms[0];
  out->ms_dedup = h->last_ms[1];
  out->ms_collect = h->last_ms[2];
  out->ms_copy = h->last_ms[3];
  out->changed_chunks = h->ctl_h ? h->ctl_h->changed : 0;
  out->ms_leaf = h->last_ms[4];
  out->leaf_labels_written = h->last_leaf_labels;
  out->leaf_digests_stored = h->last_leaf_digests;
  return MTC_OK;
}

int mtc_restart(const void* const* diffs, const uint64_t* sizes, uint32_t count, uint32_t chkpt_id,
                int diffs_on_device, void* dev_out, uint64_t out_len, int device, void* stream) {
  return guarded(nullptr, [&] {
    if (!dev_out) throw Status{MTC_ERR_INVALID_ARG, "no output buffer"};
    use_device(device);
    mtc::restart_impl(diffs, sizes, count, chkpt_id, diffs_on_device != 0, static_cast<uint8_t*>(dev_out), out_len,
                      static_cast<hipStream_t>(stream), false);
  });
}

int mtc_restart_basic(const void* const* diffs, const uint64_t* sizes, uint32_t count, uint32_t chkpt_id,
                int diffs_on_device, void* dev_out, uint64_t out_len, int device, void* stream) {
  return guarded(nullptr, [&] {
    if (!dev_out) throw Status{MTC_ERR_INVALID_ARG, "no output buffer"};
    use_device(device);
    mtc::restart_impl(diffs, sizes, count, chkpt_id, diffs_on_device != 0, static_cast<uint8_t*>(dev_out), out_len,
                      static_cast<hipStream_t>(stream), true);
  });
}

int mtc_restart_noalloc(const void* const* dev_diffs, const uint64_t* sizes, uint32_t count, uint32_t chkpt_id,
                        void* dev_out, uint64_t out_len, int device, void* stream) {
  return guarded(nullptr, [&] {
    if (!dev_out) throw Status{MTC_ERR_INVALID_ARG, "no output buffer"};
    use_device(device);
    mtc::restart_noalloc_impl(dev_diffs, sizes, count, chkpt_id, static_cast<uint8_t*>(dev_out), out_len,
                              static_cast<hipStream_t>(stream));
  });
}

int mtc_restart_times(float* upload_ms, float* total_ms) {
  mtc::last_restart_times(upload_ms, total_ms);
  return MTC_OK;
}

int mtc_digest_host(const void* host_data, uint64_t len, int hash_kind, uint8_t out[16]) {
  if ((!host_data && len) || !out) return MTC_ERR_INVALID_ARG;
  const uint8_t* p = static_cast<const uint8_t*>(host_data);
  mtc::Digest d = (hash_kind == MTC_HASH_MD5) ? mtc::md5_bytes(p, len) : mtc::m3_bytes(p, len);
  std::memcpy(out, &d, 16);
  return MTC_OK;
}

int mtc_hash_chunks(const void* dev_data, uint64_t len, uint32_t chunk_size, int hash_kind, void* dev_digests,
                    void* stream) {
  return guarded(nullptr, [&] {
    if (!dev_data || !dev_digests || !len || !chunk_size) throw Status{MTC_ERR_INVALID_ARG, "bad arguments"};
    if ((len + chunk_size - 1) / chunk_size > (1ull << 31)) throw Status{MTC_ERR_INVALID_ARG, "too many chunks"};
    hipPointerAttribute_t attr;
    MTC_HIP(hipPointerGetAttributes(&attr, dev_data));
    MTC_HIP(hipSetDevice(attr.device));
    mtc::hash_chunks(static_cast<const uint8_t*>(dev_data), len, chunk_size, hash_kind,
                     static_cast<mtc::Digest*>(dev_digests), static_cast<hipStream_t>(stream));
    MTC_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
  });
}

int mtc_set_compat(mtc_dedup* h, int on) {
  if (!h) return MTC_ERR_INVALID_ARG;
  h->compat = on ? 1 : 0;
  return MTC_OK;
}

int mtc_set_debug(mtc_dedup* h, int on) {
  if (!h) return MTC_ERR_INVALID_ARG;
  h->debug = on ? 1 : 0;
  return MTC_OK;
}

int mtc_copy_tree(const mtc_dedup* h, void* host_dst, uint64_t capacity) {
  if (!h || !host_dst) return MTC_ERR_INVALID_ARG;
  const uint64_t bytes = 16ull * (2ull * h->n - 1);
  if (!h->debug || !h->dbg_tree.p || !h->n) return MTC_ERR_STATE;
  if (capacity < bytes) return MTC_ERR_BUFFER_TOO_SMALL;
  return guarded(const_cast<mtc_dedup*>(h), [&] {
    MTC_HIP(hipSetDevice(h->device));
    MTC_HIP(hipMemcpy(host_dst, h->dbg_tree.p, bytes, hipMemcpyDeviceToHost));
  });
}

int mtc_copy_labels(const mtc_dedup* h, void* host_dst, uint64_t capacity) {
  if (!h || !host_dst) return MTC_ERR_INVALID_ARG;
  const uint64_t bytes = 2ull * h->n - 1;
  if (!h->debug || !h->dbg_labels.p || !h->n) return MTC_ERR_STATE;
  if (capacity < bytes) return MTC_ERR_BUFFER_TOO_SMALL;
  return guarded(const_cast<mtc_dedup*>(h), [&] {
    MTC_HIP(hipSetDevice(h->device));
    MTC_HIP(hipMemcpy(host_dst, h->dbg_labels.p, bytes, hipMemcpyDeviceToHost));
  });
}

int mtc_set_tuning(mtc_dedup* h, uint32_t leaf_tiles_per_block, float load_factor, uint32_t fuse_levels) {
  if (!h) return MTC_ERR_INVALID_ARG;
  return guarded(h, [&] {
    if (leaf_tiles_per_block == 0 || leaf_tiles_per_block > 4096 || !(load_factor >= 0.05f && load_factor <= 0.7f) ||
        fuse_levels == 0 || fuse_levels > 8)
      throw Status{MTC_ERR_INVALID_ARG, "tuning out of range (tiles 1..4096, load factor [0.05, 0.7], fuse 1..8)"};
    h->tune.leaf_tiles_per_block = leaf_tiles_per_block;
    h->tune.fuse = fuse_levels;
    if (h->tune.load_factor != load_factor || h->has_state) {
      // Load factor and fuse depth shape the table and the tree plan: they take
      // effect at the next baseline.
      h->tune.load_factor = load_factor;
    }
  });
}

int mtc_set_sweep_tiles(mtc_dedup* h, uint32_t tiles_a, uint32_t tiles_b) {
  if (!h) return MTC_ERR_INVALID_ARG;
  return guarded(h, [&] {
    if (tiles_a > 64 || tiles_b > 64) throw Status{MTC_ERR_INVALID_ARG, "sweep tiles per wavefront out of range (0..64)"};
    h->tune.sweep_tiles_a = tiles_a;
    h->tune.sweep_tiles_b = tiles_b;
  });
}

int mtc_batch_create(mtc_batch** out, int device, int use_graph) {
  if (!out) return MTC_ERR_INVALID_ARG;
  *out = nullptr;
  mtc_batch* b = nullptr;
  int rc = guarded(nullptr, [&] {
    use_device(device);
    b = new mtc_batch();
    b->device = device;
    b->use_graph = use_graph ? 1 : 0;
    MTC_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    MTC_HIP(hipEventCreateWithFlags(&b->ready, hipEventDisableTiming));
  });
  if (rc != MTC_OK) {
    if (b) mtc_batch_destroy(b);
    return rc;
  }
  *out = b;
  return MTC_OK;
}

int mtc_batch_destroy(mtc_batch* b) {
  if (!b) return MTC_OK;
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  b->step.reset_graph();
  if (b->ready) (void)hipEventDestroy(b->ready);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return MTC_OK;
}

int mtc_checkpoint_batch(mtc_batch* b, mtc_dedup* const* handles, const void* const* dev_data, const uint64_t* lens,
                         const int* make_baseline, uint32_t count, int advance_id, void* stream,
                         mtc_header* headers_out, uint64_t* diff_sizes_out) {
  if (!b || (count && (!handles || !dev_data || !lens))) return MTC_ERR_INVALID_ARG;
  int rc = guarded(nullptr, [&] {
    if (count == 0) return;
    if (count > mtc::kMaxBatch) throw Status{MTC_ERR_INVALID_ARG, "more than 65536 members in one batch"};
    MTC_HIP(hipSetDevice(b->device));
    for (uint32_t i = 0; i < count; i++) {
      if (!handles[i] || !dev_data[i] || lens[i] == 0) throw Status{MTC_ERR_INVALID_ARG, "empty member in batch"};
      if (handles[i]->device != b->device) throw Status{MTC_ERR_INVALID_ARG, "member on another device"};
    }
    // Work the caller queued on `stream` (the inputs) comes first.
    MTC_HIP(hipEventRecord(b->ready, static_cast<hipStream_t>(stream)));
    MTC_HIP(hipStreamWaitEvent(b->stream, b->ready, 0));
    mtc::checkpoint_batch_impl(handles, reinterpret_cast<const uint8_t* const*>(dev_data), lens, make_baseline, count,
                               b->step, b->stream, b->use_graph != 0);
    for (uint32_t i = 0; i < count; i++) {
      if (headers_out) headers_out[i] = handles[i]->last_header;
      if (diff_sizes_out) diff_sizes_out[i] = handles[i]->last_diff_size;
      if (advance_id) handles[i]->current_id += 1;
    }
  });
  if (rc != MTC_OK) b->last_error = g_last_error;
  return rc;
}

const char* mtc_batch_last_error(const mtc_batch* b) { return b ? b->last_error.c_str() : g_last_error.c_str(); }

const char* mtc_last_error(const mtc_dedup* h) { return h ? h->last_error.c_str() : g_last_error.c_str(); }

const char* mtc_status_string(int status) {
  switch (status) {
    case MTC_OK: return "ok";
    case MTC_ERR_INVALID_ARG: return "invalid argument";
    case MTC_ERR_UNSUPPORTED_APPROACH: return "unsupported approach";
    case MTC_ERR_OUT_OF_MEMORY: return "out of device memory";
    case MTC_ERR_TABLE_FULL: return "digest table full";
    case MTC_ERR_CORRUPT_DIFF: return "corrupt diff";
    case MTC_ERR_HIP: return "HIP runtime error";
    case MTC_ERR_STATE: return "invalid state";
    case MTC_ERR_BUFFER_TOO_SMALL: return "buffer too small";
    case MTC_PENDING: return "pending";
    default: return "unknown status";
  }
}

}  // extern "C"
