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

}  // namespace ps
