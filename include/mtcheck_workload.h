/*
 * mtcheck_workload.h -- synthetic checkpoint-series generator kernels
 * (libmtcheck_amd.so).  Counterpart of the reference's data generator
 * (include/data_generation.hpp:35-95, src/data_generation.cpp), which mutates a
 * buffer between checkpoints; here the mutation plan (runs and copies) is drawn
 * on the host and applied on the device, so benchmarks start with inputs resident
 * in HBM.  Not part of the deduplication path itself.
 *
 * Byte b of the pseudo-random stream with key K is byte (b & 7), little-endian, of
 * splitmix64_mix(K + ((b >> 3) + 1) * 0x9E3779B97F4A7C15).
 */
#ifndef MTCHECK_WORKLOAD_H
#define MTCHECK_WORKLOAD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Overwrite bytes [begin, end) of a device buffer with the stream of `key`. */
int mtc_wl_fill_stream(void* dev_buf, uint64_t begin, uint64_t end, uint64_t key, void* stream);

/* runs: device array of (first chunk, chunk count) uint32 pairs.  zero != 0 fills
 * the runs with zeros, otherwise with the stream of `key` at their own offsets.
 * Chunks are clipped to len. */
int mtc_wl_fill_runs(void* dev_buf, uint64_t len, uint32_t chunk_size, const uint32_t* dev_runs, uint32_t nruns,
                     uint64_t key, int zero, void* stream);

/* pairs: device array of (destination chunk, source chunk) uint32 pairs with
 * distinct destinations.  Every source is read before any destination is written;
 * dev_tmp must hold npairs * chunk_size bytes.  A short last chunk reads as
 * zero-padded and is written only up to len. */
int mtc_wl_copy_chunks(void* dev_buf, uint64_t len, uint32_t chunk_size, const uint32_t* dev_pairs, uint32_t npairs,
                       void* dev_tmp, void* stream);

#ifdef __cplusplus
}
#endif
#endif
