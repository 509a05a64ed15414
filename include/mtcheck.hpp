// mtcheck.hpp -- header-only C++ TreeDeduplicator over the C ABI of mtcheck.h.
//
// Mirrors the reference class (include/tree_approach.hpp:16-205): the same four
// checkpoint overloads with the same id-advance rules (src/tree_approach.cpp
// :1577-1776), three restart overloads (:1786-1840) and the two log writers
// (:1853-1915).  Kokkos views become raw device pointers and std::vector<uint8_t>
// host buffers; failures throw mtcheck::Error carrying the mtc_status.
#ifndef MTCHECK_HPP
#define MTCHECK_HPP

#include <chrono>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "mtcheck.h"

namespace mtcheck {

class Error : public std::runtime_error {
 public:
  Error(int status, const std::string& what) : std::runtime_error(what), status_(status) {}
  int status() const { return status_; }

 private:
  int status_;
};

inline void check(int st, const mtc_dedup* h = nullptr) {
  if (st != MTC_OK) throw Error(st, std::string(mtc_status_string(st)) + ": " + mtc_last_error(h));
}

// Mode flags, src/utils.cpp:15-36.
inline mtc_approach get_mode(int argc, char** argv) {
  static const std::pair<const char*, mtc_approach> kModes[] = {
      {"--run-full-chkpt", MTC_APPROACH_FULL},
      {"--run-basic-chkpt", MTC_APPROACH_BASIC},
      {"--run-list-chkpt", MTC_APPROACH_LIST},
      {"--run-tree-chkpt", MTC_APPROACH_TREE},
      {"--run-tree-low-offset-ref-chkpt", MTC_APPROACH_TREE_LOW_OFFSET_REF},
      {"--run-tree-low-offset-chkpt", MTC_APPROACH_TREE_LOW_OFFSET},
      {"--run-tree-low-root-ref-chkpt", MTC_APPROACH_TREE_LOW_ROOT_REF},
      {"--run-tree-low-root-chkpt", MTC_APPROACH_TREE_LOW_ROOT},
  };
  for (int i = 0; i < argc; i++)
    for (const auto& m : kModes)
      if (std::strcmp(argv[i], m.first) == 0) return m.second;
  return MTC_APPROACH_UNKNOWN;
}

// calculate_digest_host, include/hash_functions.hpp:46-82: hex in groups of four bytes.
inline std::string calculate_digest_host(const void* data, uint64_t len, int hash_kind = MTC_HASH_MURMUR3) {
  uint8_t d[16];
  check(mtc_digest_host(data, len, hash_kind, d));
  static const char* hex = "0123456789ABCDEF";
  std::string s;
  for (int i = 0; i < 16; i++) {
    s += hex[d[i] >> 4];
    s += hex[d[i] & 15];
    if (i % 4 == 3) s += ' ';
  }
  return s;
}

// write_metadata_breakdown, src/utils.cpp:38-124 (tree mode: region sizes do not
// enter the line; the repeat map does).
inline void write_metadata_breakdown(std::ostream& fs, const mtc_header& h, const std::vector<uint8_t>& diff,
                                     uint32_t num_chkpts) {
  fs << sizeof(mtc_header) << "," << h.num_first_ocur * 4ull << ",";
  if (h.ref_id != h.chkpt_id) {
    fs << 8ull * h.num_prior_chkpts;
    for (uint32_t i = 0; i < num_chkpts; i++) {
      uint64_t bytes = 0;
      if (i < h.num_prior_chkpts) {
        uint32_t num = 0;
        std::memcpy(&num, diff.data() + sizeof(mtc_header) + 4ull * h.num_first_ocur + 8ull * i + 4, 4);
        bytes = 8ull * num;
      }
      fs << "," << bytes;
    }
  } else {
    fs << 0 << "," << 8ull * h.num_shift_dupl;
    for (uint32_t i = 1; i < num_chkpts; i++) fs << "," << 0;
  }
  fs << "\n";
}

class TreeDeduplicator {
 public:
  explicit TreeDeduplicator(uint32_t bytes_per_chunk, int hash_kind = MTC_HASH_MURMUR3, int device = 0,
                            int approach = MTC_APPROACH_TREE)
      : chunk_size(bytes_per_chunk), device_(device), approach_(approach) {
    check(mtc_create(&h_, bytes_per_chunk, approach, hash_kind, device));
  }
  ~TreeDeduplicator() { mtc_destroy(h_); }
  TreeDeduplicator(const TreeDeduplicator&) = delete;
  TreeDeduplicator& operator=(const TreeDeduplicator&) = delete;

  // mtc_set_compat: the reference's bytes where this library departs from them.
  void set_compat(bool on) { check(mtc_set_compat(h_, on ? 1 : 0), h_); }
  size_t num_first_ocur() const { return stats().num_first_ocur; }  // include/tree_approach.hpp:53-55
  size_t num_shift_dupl() const { return stats().num_shift_dupl; }  // include/tree_approach.hpp:57-59
  uint32_t current_id() const { return stats().current_id; }
  mtc_stats stats() const {
    mtc_stats s;
    check(mtc_get_stats(h_, &s), h_);
    return s;
  }
  mtc_dedup* handle() const { return h_; }

  // checkpoint(header_t&, uint8_t*, size_t, HostMirror&, bool), :1577-1689: the id does not advance.
  void checkpoint(mtc_header& header, const uint8_t* data_ptr, size_t data_len, std::vector<uint8_t>& diff_h,
                  bool make_baseline, void* stream = nullptr) {
    run(header, data_ptr, data_len, diff_h, make_baseline, false, stream);
  }
  // checkpoint(ptr, len, filename, logname, bool), :1705-1724: diff file + logs, id advances.
  void checkpoint(const uint8_t* data_ptr, size_t len, const std::string& filename, const std::string& logname,
                  bool make_baseline, void* stream = nullptr) {
    mtc_header h;
    std::vector<uint8_t> diff;
    run(h, data_ptr, len, diff, make_baseline, true, stream);
    write_chkpt_log(h, diff, logname);
    std::ofstream f(filename, std::ios::binary | std::ios::trunc);
    f.write(reinterpret_cast<const char*>(diff.data()), (std::streamsize)diff.size());
    if (!f) throw Error(MTC_ERR_INVALID_ARG, "cannot write " + filename);
  }
  // checkpoint(ptr, len, diff_h, bool), :1740-1748: id advances.
  void checkpoint(const uint8_t* data_ptr, size_t len, std::vector<uint8_t>& diff_h, bool make_baseline,
                  void* stream = nullptr) {
    mtc_header h;
    run(h, data_ptr, len, diff_h, make_baseline, true, stream);
  }
  // checkpoint(ptr, len, diff_h, logname, bool), :1765-1776: logs, id advances.
  void checkpoint(const uint8_t* data_ptr, size_t len, std::vector<uint8_t>& diff_h, const std::string& logname,
                  bool make_baseline, void* stream = nullptr) {
    mtc_header h;
    run(h, data_ptr, len, diff_h, make_baseline, true, stream);
    write_chkpt_log(h, diff_h, logname);
  }

  // restart(View, vector<HostMirror>&, logname, id), :1786-1797 and the raw-pointer
  // overload :1809-1817: host diffs, output into device memory.
  void restart(uint8_t* data_ptr, size_t len, const std::vector<std::vector<uint8_t>>& chkpts,
               const std::string& logname, uint32_t chkpt_id, void* stream = nullptr) {
    std::vector<const void*> ptrs;
    std::vector<uint64_t> sizes;
    for (const auto& c : chkpts) {
      ptrs.push_back(c.empty() ? nullptr : c.data());
      sizes.push_back(c.size());
    }
    run_restart(ptrs, sizes, 0, data_ptr, len, chkpt_id, stream);
    if (!logname.empty()) write_restart_log(chkpt_id, logname + ".chunk_size." + std::to_string(chunk_size) +
                                                          ".restart_timing.csv");
  }
  // Diffs already in device memory (pointer, size per checkpoint id).
  void restart(uint8_t* data_ptr, size_t len, const std::vector<std::pair<const void*, uint64_t>>& dev_chkpts,
               const std::string& logname, uint32_t chkpt_id, void* stream = nullptr) {
    std::vector<const void*> ptrs;
    std::vector<uint64_t> sizes;
    for (const auto& c : dev_chkpts) {
      ptrs.push_back(c.first);
      sizes.push_back(c.second);
    }
    run_restart(ptrs, sizes, 1, data_ptr, len, chkpt_id, stream);
    if (!logname.empty()) write_restart_log(chkpt_id, logname);
  }
  // restart(View, vector<string>& names, logname, id), :1827-1840: reads
  // <name>.hashtree.incr_chkpt for every name.
  void restart(uint8_t* data_ptr, size_t len, const std::vector<std::string>& chkpt_filenames,
               const std::string& logname, uint32_t chkpt_id, void* stream = nullptr) {
    std::vector<std::vector<uint8_t>> diffs;
    for (const auto& name : chkpt_filenames) {
      std::ifstream f(name + ".hashtree.incr_chkpt", std::ios::binary | std::ios::ate);
      if (!f) throw Error(MTC_ERR_INVALID_ARG, "cannot open " + name + ".hashtree.incr_chkpt");
      std::vector<uint8_t> d((size_t)f.tellg());
      f.seekg(0);
      f.read(reinterpret_cast<char*>(d.data()), (std::streamsize)d.size());
      diffs.push_back(std::move(d));
    }
    std::vector<const void*> ptrs;
    std::vector<uint64_t> sizes;
    for (const auto& c : diffs) {
      ptrs.push_back(c.data());
      sizes.push_back(c.size());
    }
    run_restart(ptrs, sizes, 0, data_ptr, len, chkpt_id, stream);
    if (!logname.empty()) write_restart_log(chkpt_id, logname);
  }

  // write_chkpt_log, :1853-1896 (basic: src/basic_approach.cpp:632-673): <logname>.chunk_size.<C>.csv,
  // .size.csv, .timing.csv.
  void write_chkpt_log(const mtc_header& header, const std::vector<uint8_t>& diff_h, const std::string& logname) {
    const std::string base = logname + ".chunk_size." + std::to_string(chunk_size);
    std::ofstream result(base + ".csv", std::ios::app), size(base + ".size.csv", std::ios::app),
        timing(base + ".timing.csv", std::ios::app);
    if (result.tellp() == 0)
      result << "Approach,Chkpt ID,Chunk Size,Uncompressed Size,Compressed Size,Data Size,Metadata Size,"
                "Setup Time,Comparison Time,Gather Time,Write Time\n";
    const uint32_t cid = header.chkpt_id;
    const char* name = approach_name();
    result << name << "," << cid << "," << chunk_size << "," << header.datalen << "," << datasizes.first + datasizes.second
           << "," << datasizes.first << "," << datasizes.second << "," << timers[0] << "," << timers[1] << ","
           << timers[2] << "," << timers[3] << "\n";
    timing << name << "," << cid << "," << chunk_size << "," << timers[1] << "," << timers[2] << "," << timers[3] << "\n";
    size << name << "," << cid << "," << chunk_size << "," << datasizes.first << "," << datasizes.second << ",";
    // the basic log counts the checkpoints since the baseline (:672), the tree log the id (:1895)
    write_metadata_breakdown(size, header, diff_h, approach_ == MTC_APPROACH_BASIC ? cid - header.ref_id + 1 : cid);
  }
  // write_restart_log, :1904-1915 (basic: src/basic_approach.cpp:682-691).
  void write_restart_log(uint32_t select_chkpt, const std::string& logname) {
    std::ofstream f(logname, std::ios::app);
    f << approach_name() << "," << select_chkpt << "," << chunk_size << "," << restart_timers[0] << ","
      << restart_timers[1] << "\n";
  }
  const char* approach_name() const { return approach_ == MTC_APPROACH_BASIC ? "Basic" : "Tree"; }

  const uint32_t chunk_size;
  std::pair<uint64_t, uint64_t> datasizes{0, 0};  // (data incl. header, metadata), :707-708
  double timers[4] = {0, 0, 0, 0};                // seconds, :1640,1657,1672,1688
  double restart_timers[2] = {0, 0};

 private:
  void run(mtc_header& header, const uint8_t* data_ptr, size_t len, std::vector<uint8_t>& diff, bool make_baseline,
           bool advance, void* stream) {
    uint64_t size = 0;
    check(mtc_checkpoint(h_, data_ptr, len, make_baseline, advance, stream, &header, &size), h_);
    diff.resize(size);
    // the copy of the diff to the host is the fourth timer (:1678-1688)
    const auto t0 = std::chrono::steady_clock::now();
    check(mtc_copy_diff(h_, diff.data(), size, stream), h_);
    const double copy_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    const mtc_stats s = stats();
    const uint64_t meta = 4ull * header.num_first_ocur + 8ull * header.num_prior_chkpts + 8ull * header.num_shift_dupl;
    datasizes = {size - meta, meta};
    timers[0] = s.ms_setup / 1e3;
    timers[1] = s.ms_dedup / 1e3;
    timers[2] = s.ms_collect / 1e3;
    timers[3] = copy_s;
  }
  void run_restart(const std::vector<const void*>& ptrs, const std::vector<uint64_t>& sizes, int on_device,
                   uint8_t* out, size_t len, uint32_t chkpt_id, void* stream) {
    auto fn = approach_ == MTC_APPROACH_BASIC ? mtc_restart_basic : mtc_restart;
    check(fn(ptrs.data(), sizes.data(), (uint32_t)ptrs.size(), chkpt_id, on_device, out, len, device_, stream));
    float up = 0, total = 0;
    mtc_restart_times(&up, &total);
    restart_timers[0] = up / 1e3;
    restart_timers[1] = total / 1e3;
  }

  mtc_dedup* h_ = nullptr;
  int device_;
  int approach_;
};

// BasicDeduplicator's restart overloads (src/basic_approach.cpp:593-630) and its
// restart log (:681-690) over mtc_restart_basic: restores chains of basic diffs,
// such as the .basic.incr_chkpt files the reference writes.  It holds no
// deduplicator handle (basic checkpoints are taken by a TreeDeduplicator created
// with MTC_APPROACH_BASIC, whose restart overloads call mtc_restart_basic too).
class BasicRestarter {
 public:
  explicit BasicRestarter(uint32_t chunk_size, int device = 0) : chunk_size(chunk_size), device_(device) {}

  // restart(View, vector<HostMirror>&, logname, id), :593-614.
  void restart(uint8_t* data_ptr, size_t len, const std::vector<std::vector<uint8_t>>& chkpts,
               const std::string& logname, uint32_t chkpt_id, void* stream = nullptr) {
    run(chkpts, data_ptr, len, chkpt_id, stream);
    if (!logname.empty()) write_restart_log(chkpt_id, logname + ".chunk_size." + std::to_string(chunk_size) +
                                                          ".restart_timing.csv");
  }
  // restart(View, vector<string>& names, logname, id), :616-630: reads
  // <name>.basic.incr_chkpt for every name.
  void restart(uint8_t* data_ptr, size_t len, const std::vector<std::string>& chkpt_filenames,
               const std::string& logname, uint32_t chkpt_id, void* stream = nullptr) {
    std::vector<std::vector<uint8_t>> diffs;
    for (const auto& name : chkpt_filenames) {
      std::ifstream f(name + ".basic.incr_chkpt", std::ios::binary | std::ios::ate);
      if (!f) throw Error(MTC_ERR_INVALID_ARG, "cannot open " + name + ".basic.incr_chkpt");
      std::vector<uint8_t> d((size_t)f.tellg());
      f.seekg(0);
      f.read(reinterpret_cast<char*>(d.data()), (std::streamsize)d.size());
      diffs.push_back(std::move(d));
    }
    run(diffs, data_ptr, len, chkpt_id, stream);
    if (!logname.empty()) write_restart_log(chkpt_id, logname);
  }
  // write_restart_log, :681-690.
  void write_restart_log(uint32_t select_chkpt, const std::string& logname) {
    std::ofstream f(logname, std::ios::app);
    f << "Basic," << select_chkpt << "," << chunk_size << "," << restart_timers[0] << "," << restart_timers[1] << "\n";
  }

  const uint32_t chunk_size;
  double restart_timers[2] = {0, 0};

 private:
  void run(const std::vector<std::vector<uint8_t>>& chkpts, uint8_t* out, size_t len, uint32_t chkpt_id,
           void* stream) {
    std::vector<const void*> ptrs;
    std::vector<uint64_t> sizes;
    for (const auto& c : chkpts) {
      ptrs.push_back(c.empty() ? nullptr : c.data());
      sizes.push_back(c.size());
    }
    check(mtc_restart_basic(ptrs.data(), sizes.data(), (uint32_t)ptrs.size(), chkpt_id, 0, out, len, device_,
                            stream));
    float up = 0, total = 0;
    mtc_restart_times(&up, &total);
    restart_timers[0] = up / 1e3;
    restart_timers[1] = total / 1e3;
  }

  int device_;
};

}  // namespace mtcheck

#endif
