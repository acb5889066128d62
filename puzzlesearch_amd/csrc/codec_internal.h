// Chunk encoders shared by the codec entry points and the segment step.
// `starts` (host) delimits the chunks of d_values; `off` receives the byte
// offset of each encoded chunk in d_out (nchunks + 1 entries).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace ps {

constexpr uint32_t MAX_CHUNK_VALUES = 256 * 1024;  // StreamVInt::MAX_INDEXES_COUNT

// StreamVInt::Encode, with bounds when d_bounds is given (Common/StreamVInt.cpp:209-279)
int svi_encode_chunks(const uint32_t* d_values, const uint8_t* d_bounds, const std::vector<uint64_t>& starts,
                      uint8_t* d_out, uint64_t out_cap, std::vector<uint64_t>& off, hipStream_t s);
// FrontierCompression::EncodeWithCheck (with_check) or Encode (Common/FrontierCompression.cpp:92-115);
// d_offsets (optional, device) receives a copy of off
int fc_encode_chunks(const uint32_t* d_values, const std::vector<uint64_t>& starts, int with_check, uint8_t* d_out,
                     uint64_t out_cap, std::vector<uint64_t>& off, uint64_t* d_offsets, hipStream_t s);

// Stream-ordered variants: every per-chunk decision and offset is made on the
// device, d_offsets (device, nchunks + 1) receives the chunk offsets and
// nothing waits for the stream.  The caller keeps d_out large enough for the
// worst case and reads the offsets back once the stream is done.
// fc_encode_chunks_dev counts unsorted unchecked bitmap chunks in *d_err (the
// host-planned encoder refuses them); bitmaps wider than LDS go through a
// global scratch.
// svi_encode_chunks_dev writes no chunk that would end past out_cap and
// counts such chunks in *d_err (optional).
int svi_encode_chunks_dev(const uint32_t* d_values, const uint8_t* d_bounds, const std::vector<uint64_t>& starts,
                          uint8_t* d_out, uint64_t out_cap, uint64_t* d_offsets, uint32_t* d_err, hipStream_t s);
int fc_encode_chunks_dev(const uint32_t* d_values, const std::vector<uint64_t>& starts, int with_check,
                         uint8_t* d_out, uint64_t* d_offsets, uint32_t* d_err, hipStream_t s);

}  // namespace ps
