// restart_chkpt_files -- rebuild a checkpoint from its chain of diffs.
//
// Usage (src/restart_chkpt_files.cpp:21-46):
//   restart_chkpt_files restart_id num_chkpts num_tests chunk_size <mode> file0 file1 ...
// The original files give the restored size; the diffs are <file>.hashtree.incr_chkpt
// for the tree modes and <file>.basic.incr_chkpt for --run-basic-chkpt (:109-136).
// Prints the digest of the restored buffer, as the reference's VERIFY_OUTPUT build
// does (:192-194), and the digest of the original file beside it.
#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "mtcheck.hpp"

#define HIP_OK(x)                                                              \
  do {                                                                         \
    hipError_t e_ = (x);                                                       \
    if (e_ != hipSuccess) {                                                    \
      std::fprintf(stderr, "HIP error %s at %s:%d\n", hipGetErrorString(e_), __FILE__, __LINE__); \
      return 3;                                                                \
    }                                                                          \
  } while (0)

static std::vector<uint8_t> read_file(const std::string& path) {
  std::ifstream f(path, std::ios::binary | std::ios::ate);
  if (!f) throw mtcheck::Error(MTC_ERR_INVALID_ARG, "cannot open " + path);
  std::vector<uint8_t> b((size_t)f.tellg());
  f.seekg(0);
  f.read(reinterpret_cast<char*>(b.data()), (std::streamsize)b.size());
  return b;
}

int main(int argc, char** argv) {
  if (argc < 6) {
    std::fprintf(stderr, "usage: %s restart_id num_chkpts num_tests chunk_size <mode> file0 [file1 ...]\n", argv[0]);
    return 1;
  }
  const uint32_t restart_id = (uint32_t)std::atoi(argv[1]);
  const uint32_t num_chkpts = (uint32_t)std::atoi(argv[2]);
  const uint32_t num_tests = (uint32_t)std::atoi(argv[3]);
  const uint32_t chunk_size = (uint32_t)std::atoi(argv[4]);
  const mtc_approach mode = mtcheck::get_mode(argc, argv);
  if (mode == MTC_APPROACH_FULL || mode == MTC_APPROACH_LIST) {
    std::fprintf(stderr, "ERROR: only the tree and basic approaches are provided\n");
    return 2;
  }
  if ((uint32_t)argc < 6 + num_chkpts || restart_id >= num_chkpts) {
    std::fprintf(stderr, "ERROR: bad checkpoint count or restart id\n");
    return 1;
  }
  try {
    std::vector<std::string> files, trimmed;
    for (uint32_t i = 0; i < num_chkpts; i++) {
      files.push_back(argv[6 + i]);
      trimmed.push_back(files.back().substr(files.back().rfind('/') + 1));
    }
    const std::vector<uint8_t> original = read_file(files[restart_id]);
    std::vector<std::vector<uint8_t>> chkpts;
    const bool basic = mode == MTC_APPROACH_BASIC;
    for (const auto& f : files) chkpts.push_back(read_file(f + (basic ? ".basic.incr_chkpt" : ".hashtree.incr_chkpt")));
    uint8_t* dev = nullptr;
    HIP_OK(hipMalloc(&dev, original.size() ? original.size() : 1));
    std::vector<uint8_t> restored(original.size());
    for (uint32_t t = 0; t < num_tests; t++) {
      HIP_OK(hipMemset(dev, 0, original.size()));
      if (basic) {
        mtcheck::BasicRestarter dedup(chunk_size);
        dedup.restart(dev, original.size(), chkpts, trimmed[restart_id], restart_id);
      } else {
        mtcheck::TreeDeduplicator dedup(chunk_size);
        dedup.restart(dev, original.size(), chkpts, trimmed[restart_id], restart_id);
      }
      HIP_OK(hipMemcpy(restored.data(), dev, original.size(), hipMemcpyDeviceToHost));
      std::cout << (basic ? "Basic chkpt digest:    " : "Hashtree digest:       ")
                << mtcheck::calculate_digest_host(restored.data(), restored.size()) << std::endl;
    }
    std::cout << "Original digest:       " << mtcheck::calculate_digest_host(original.data(), original.size())
              << std::endl;
    HIP_OK(hipFree(dev));
    if (restored != original) {
      std::fprintf(stderr, "ERROR: restored checkpoint differs from the original\n");
      return 5;
    }
  } catch (const mtcheck::Error& e) {
    std::fprintf(stderr, "ERROR: %s\n", e.what());
    return 4;
  }
  return 0;
}
