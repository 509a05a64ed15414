// dedup_chkpt_files -- deduplicate a series of checkpoint files on the GPU.
//
// Usage (src/dedup_files.cpp:11-24,42-63):
//   dedup_chkpt_files chunk_size num_chkpts <mode> file0 file1 ...
// Writes <file>.hashtree.incr_chkpt (--run-basic-chkpt: <file>.basic.incr_chkpt)
// next to each file and appends the three CSV logs named after the file's base
// name in the working directory (:87-100).  Every tree mode runs the tree
// deduplicator and Basic the basic one; Full and List are not part of this library
// and are rejected.
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

int main(int argc, char** argv) {
  if (argc < 4) {
    std::fprintf(stderr, "usage: %s chunk_size num_chkpts <mode> file0 [file1 ...]\n", argv[0]);
    return 1;
  }
  const uint32_t chunk_size = (uint32_t)std::atoi(argv[1]);
  const uint32_t num_chkpts = (uint32_t)std::atoi(argv[2]);
  const mtc_approach mode = mtcheck::get_mode(argc, argv);
  if (mode == MTC_APPROACH_UNKNOWN) std::fprintf(stderr, "ERROR: Incorrect mode; running the tree approach\n");
  if (mode == MTC_APPROACH_FULL || mode == MTC_APPROACH_LIST) {
    std::fprintf(stderr, "ERROR: only the tree and basic approaches are provided\n");
    return 2;
  }
  const bool basic = mode == MTC_APPROACH_BASIC;
  if ((uint32_t)argc < 4 + num_chkpts) {
    std::fprintf(stderr, "ERROR: %u checkpoint files expected\n", num_chkpts);
    return 1;
  }
  try {
    mtcheck::TreeDeduplicator dedup(chunk_size, MTC_HASH_MURMUR3, 0, basic ? MTC_APPROACH_BASIC : MTC_APPROACH_TREE);
    uint8_t* dev = nullptr;
    size_t dev_len = 0;
    for (uint32_t idx = 0; idx < num_chkpts; idx++) {
      const std::string path = argv[4 + idx];
      std::ifstream f(path, std::ios::binary | std::ios::ate);
      if (!f) throw mtcheck::Error(MTC_ERR_INVALID_ARG, "cannot open " + path);
      const size_t len = (size_t)f.tellg();
      f.seekg(0);
      std::vector<uint8_t> host(len);
      f.read(reinterpret_cast<char*>(host.data()), (std::streamsize)len);
      if (len > dev_len) {
        if (dev) HIP_OK(hipFree(dev));
        HIP_OK(hipMalloc(&dev, len));
        dev_len = len;
      }
      HIP_OK(hipMemcpy(dev, host.data(), len, hipMemcpyHostToDevice));
      std::string logname = path.substr(path.rfind('/') + 1);
      dedup.checkpoint(dev, len, path + (basic ? ".basic.incr_chkpt" : ".hashtree.incr_chkpt"), logname, idx == 0);
      std::printf("checkpoint %u: %s first=%zu shift=%zu\n", idx, path.c_str(), dedup.num_first_ocur(),
                  dedup.num_shift_dupl());
    }
    if (dev) HIP_OK(hipFree(dev));
  } catch (const mtcheck::Error& e) {
    std::fprintf(stderr, "ERROR: %s\n", e.what());
    return 4;
  }
  return 0;
}
