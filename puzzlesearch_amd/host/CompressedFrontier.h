// Compressed segment streams over a Store, with the interface of the
// reference's Common/CompressedFrontier.h:60-124 and the same bytes: a
// segment's stream is a sequence of FrontierCompression chunks (a bitmap chunk
// `u32 words | 0x80000000, u32 first, words x u64`, or a StreamVInt chunk `u32
// count` + 4-value tuples), each of at most 2^18 values, written to the Store
// in pieces of whole chunks.  Encoding and decoding run on the device
// (ps_fc_encode / ps_fc_decode); the host only finds chunk boundaries.
#pragma once

#include <cstdint>
#include <vector>

#include "Buffer.h"
#include "Store.h"
#include "puzzlesearch_amd.h"

namespace fcstream {

constexpr size_t MAX_CHUNK_VALUES = 1u << 18;         // StreamVInt::MAX_INDEXES_COUNT
constexpr size_t FRONTIER_PIECE = 12 * 1024 * 1024;  // CompressedSegmentWriter's output buffer
constexpr size_t XSEG_PIECE = 16 * 1024 * 1024;      // CompressedCrossSegmentReader's input buffer

// Byte offsets of the chunks in buf[0, size) (offsets.back() == size) and
// the number of values they hold.  Throws std::runtime_error on a truncated
// or malformed stream (Common/FrontierCompression.cpp:117-125,
// Common/StreamVInt.cpp:233-251 give each chunk's length).
uint64_t ChunkBoundaries(const uint8_t* buf, size_t size, std::vector<uint64_t>& offsets);

// Writes bytes to `segment` of `store` in pieces of at most `piece` bytes,
// cut at the given chunk boundaries (chunk_offsets relative to bytes).
void WritePieces(Store& store, int segment, const uint8_t* bytes, const std::vector<uint64_t>& chunk_offsets,
                 size_t piece);

// A grow-only device buffer owned by the host code.
class DeviceBytes {
public:
    DeviceBytes() = default;
    ~DeviceBytes();
    DeviceBytes(const DeviceBytes&) = delete;
    DeviceBytes& operator=(const DeviceBytes&) = delete;
    void* Reserve(uint64_t bytes);
    void* Ptr() const { return m_Ptr; }

private:
    void* m_Ptr = nullptr;
    uint64_t m_Cap = 0;
};

// Decodes a stream of FrontierCompression chunks on the device into `ids`
// (device, grown as needed); returns the number of values.
uint64_t DecodeToDevice(const std::vector<uint8_t>& stream, DeviceBytes& ids, DeviceBytes& scratch, ps_stream s);

// Encodes n device ids as FrontierCompression chunks of at most
// MAX_CHUNK_VALUES values (EncodeWithCheck when with_check) into host bytes;
// chunk_offsets receives the chunk boundaries.
void EncodeFromDevice(const uint32_t* d_ids, uint64_t n, size_t chunk_values, bool with_check,
                      std::vector<uint8_t>& bytes, std::vector<uint64_t>& chunk_offsets, DeviceBytes& scratch,
                      ps_stream s);

}  // namespace fcstream

// Common/CompressedFrontier.h:57-75
class CompressedSegmentReader {
public:
    explicit CompressedSegmentReader(Store& store);
    void SetSegment(int segment);
    void Delete(int segment);
    // the next chunk of the segment's ids (empty at the end)
    Buffer<uint32_t>& Read();
    // every remaining id of the segment, decoded into device memory; returns their number
    uint64_t ReadAllToDevice(fcstream::DeviceBytes& ids, ps_stream s);

private:
    bool NextStoreChunk();
    Store& m_Store;
    int m_Segment = -1;
    std::vector<uint8_t> m_Input;
    std::vector<uint64_t> m_Chunks;
    size_t m_Next = 0;
    Buffer<uint32_t> m_IndexBuffer;
    fcstream::DeviceBytes m_DevOut, m_Scratch;
};

// Common/CompressedFrontier.h:77-103
class CompressedSegmentWriter {
public:
    explicit CompressedSegmentWriter(Store& store);
    void SetSegment(int segment);
    void Add(uint32_t value);
    void Flush();
    // n device ids (ascending) of the current segment, encoded on the device
    void AddFromDevice(const uint32_t* d_ids, uint64_t n, ps_stream s);

private:
    void FlushData();
    Store& m_Store;
    int m_Segment = -1;
    Buffer<uint32_t> m_IndexBuffer;
    std::vector<uint8_t> m_Output;
    fcstream::DeviceBytes m_DevIds, m_Scratch;
};

// Common/CompressedFrontier.h:105-124: the cross-segment stream of every op
class CompressedCrossSegmentReader {
public:
    explicit CompressedCrossSegmentReader(StoreSet& storeSet);
    void SetSegment(int segment);
    // the next chunk of `op`'s ids for the segment (empty at the end)
    Buffer<uint32_t>& Read(int op);
    void Delete(int segment);

private:
    StoreSet& m_StoreSet;
    int m_Segment = -1;
    int m_LastOp = -1;
    std::vector<uint8_t> m_Input;
    std::vector<uint64_t> m_Chunks;
    size_t m_Next = 0;
    Buffer<uint32_t> m_IndexBuffer;
    fcstream::DeviceBytes m_DevOut, m_Scratch;
};
