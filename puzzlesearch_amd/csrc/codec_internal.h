// Chunk encoders shared by the codec entry points and the segment steps.
// `starts` (host) delimits the chunks of d_values.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace ps {

constexpr uint32_t MAX_CHUNK_VALUES = 256 * 1024;  // StreamVInt::MAX_INDEXES_COUNT

// Every per-chunk decision and offset is made on the device, d_offsets
// (device, nchunks + 1) receives the chunk offsets and nothing waits for the
// stream: the caller reads the offsets and the error counts back once the
// stream is done.
// StreamVInt::Encode, with bounds when d_bounds is given
// (Common/StreamVInt.cpp:209-279): a chunk that would end past out_cap is not
// written and is counted in *d_err (optional).
// FrontierCompression::EncodeWithCheck (with_check) or Encode
// (Common/FrontierCompression.cpp:92-115): chunks that Encode would store as a
// bitmap but that are not sorted are counted in d_err[0], chunks that would
// end past out_cap are not written and are counted in d_err[1] (d_err
// optional); bitmaps wider than LDS go through a global scratch.
int svi_encode_chunks_dev(const uint32_t* d_values, const uint8_t* d_bounds, const std::vector<uint64_t>& starts,
                          uint8_t* d_out, uint64_t out_cap, uint64_t* d_offsets, uint32_t* d_err, hipStream_t s);
int fc_encode_chunks_dev(const uint32_t* d_values, const std::vector<uint64_t>& starts, int with_check,
                         uint8_t* d_out, uint64_t out_cap, uint64_t* d_offsets, uint32_t* d_err, hipStream_t s);

}  // namespace ps
