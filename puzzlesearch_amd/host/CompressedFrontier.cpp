// Compressed segment streams (CompressedFrontier.h) with device codecs.
#include "CompressedFrontier.h"

#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <string>

#include "Util.h"

namespace fcstream {

namespace {

uint32_t Rd32(const uint8_t* p) {
    uint32_t v;
    std::memcpy(&v, p, 4);
    return v;
}

[[noreturn]] void Malformed(size_t pos, const char* what) {
    throw std::runtime_error("malformed FrontierCompression stream at byte " + std::to_string(pos) + ": " + what);
}

}  // namespace

uint64_t ChunkBoundaries(const uint8_t* buf, size_t size, std::vector<uint64_t>& offsets) {
    offsets.assign(1, 0);
    uint64_t values = 0;
    size_t pos = 0;
    while (pos < size) {
        if (pos + 4 > size) Malformed(pos, "truncated chunk header");
        const uint32_t h = Rd32(buf + pos);
        size_t len;
        if (h & 0x80000000u) {  // bitmap: u32 words | 0x80000000, u32 first, words x u64
            const uint64_t words = h & 0x7FFFFFFFu;
            len = 8 + (size_t)words * 8;
            if (pos + len > size) Malformed(pos, "truncated bitmap chunk");
            uint64_t bits = 1;  // `first` itself
            for (uint64_t w = 0; w < words; w++) {
                uint64_t v;
                std::memcpy(&v, buf + pos + 8 + w * 8, 8);
                bits += (uint64_t)__builtin_popcountll(v);
            }
            values += bits;
        } else {  // StreamVInt: u32 count, then ceil8(count)/4 tuples of 1 + 4..16 bytes
            const uint32_t count = h;
            if (count == 0 || count > MAX_CHUNK_VALUES) Malformed(pos, "bad value count");
            const uint64_t tuples = ((uint64_t)count + 7) / 8 * 2;
            size_t p = pos + 4;
            for (uint64_t t = 0; t < tuples; t++) {
                if (p >= size) Malformed(pos, "truncated StreamVInt chunk");
                const uint32_t d = buf[p];
                p += 5 + (d & 3) + ((d >> 2) & 3) + ((d >> 4) & 3) + ((d >> 6) & 3);
            }
            if (p > size) Malformed(pos, "truncated StreamVInt chunk");
            len = p - pos;
            values += count;
        }
        pos += len;
        offsets.push_back(pos);
    }
    return values;
}

void WritePieces(Store& store, int segment, const uint8_t* bytes, const std::vector<uint64_t>& chunk_offsets,
                 size_t piece) {
    size_t start = 0;
    for (size_t c = 1; c < chunk_offsets.size(); c++) {
        const size_t end = chunk_offsets[c];
        if (end - start > piece) {  // the piece ends before chunk c-1
            const size_t cut = chunk_offsets[c - 1];
            ensure(cut > start);  // a chunk is smaller than a piece
            store.Write(segment, bytes + start, cut - start);
            start = cut;
        }
    }
    if (!chunk_offsets.empty() && chunk_offsets.back() > start)
        store.Write(segment, bytes + start, chunk_offsets.back() - start);
}

DeviceBytes::~DeviceBytes() {
    if (m_Ptr) ps_free(m_Ptr);
}

void* DeviceBytes::Reserve(uint64_t bytes) {
    if (bytes <= m_Cap && m_Ptr) return m_Ptr;
    if (m_Ptr) ps_free(m_Ptr);
    m_Ptr = nullptr;
    m_Cap = 0;
    const uint64_t want = std::max<uint64_t>(bytes + bytes / 4, 4096);
    PsCheck(ps_alloc(&m_Ptr, want), "device stream buffer");
    m_Cap = want;
    return m_Ptr;
}

uint64_t DecodeToDevice(const std::vector<uint8_t>& stream, DeviceBytes& ids, DeviceBytes& scratch, ps_stream s) {
    std::vector<uint64_t> off;
    const uint64_t n = ChunkBoundaries(stream.data(), stream.size(), off);
    if (n == 0) return 0;
    const uint64_t nchunks = off.size() - 1;
    const uint64_t in_bytes = (stream.size() + 64 + 7) / 8 * 8;  // the decoder may read past the end
    uint8_t* base = (uint8_t*)scratch.Reserve(in_bytes + (nchunks + 1) * 16);
    uint64_t* d_off = (uint64_t*)(base + in_bytes);
    uint64_t* d_vs = d_off + nchunks + 1;
    uint32_t* d_ids = (uint32_t*)ids.Reserve(n * 4 + 64);
    PsCheck(ps_memcpy_h2d_async(base, stream.data(), stream.size(), s), "stream upload");
    PsCheck(ps_memcpy_h2d_async(d_off, off.data(), off.size() * 8, s), "stream upload");
    PsCheck(ps_fc_decode(base, d_off, (uint32_t)nchunks, d_ids, n, d_vs, s), "ps_fc_decode");
    PsCheck(ps_stream_sync(s), "stream synchronize");
    return n;
}

void EncodeFromDevice(const uint32_t* d_ids, uint64_t n, size_t chunk_values, bool with_check,
                      std::vector<uint8_t>& bytes, std::vector<uint64_t>& chunk_offsets, DeviceBytes& scratch,
                      ps_stream s) {
    bytes.clear();
    chunk_offsets.assign(1, 0);
    if (n == 0) return;
    std::vector<uint64_t> starts;
    uint64_t cap = 0;
    for (uint64_t c = 0; c < n; c += chunk_values) {
        starts.push_back(c);
        const uint64_t m = std::min<uint64_t>(chunk_values, n - c);
        cap += 4 + 17 * ((m + 7) / 8 * 2);  // a bitmap is chosen only when smaller
    }
    starts.push_back(n);
    const uint64_t nchunks = starts.size() - 1;
    uint8_t* base = (uint8_t*)scratch.Reserve((nchunks + 1) * 16 + cap + 64);
    uint64_t* d_starts = (uint64_t*)base;
    uint64_t* d_off = d_starts + nchunks + 1;
    uint8_t* d_out = (uint8_t*)(d_off + nchunks + 1);
    PsCheck(ps_memcpy_h2d_async(d_starts, starts.data(), starts.size() * 8, s), "chunk starts upload");
    PsCheck(ps_fc_encode(d_ids, d_starts, (uint32_t)nchunks, with_check ? 1 : 0, d_out, cap, d_off, s), "ps_fc_encode");
    chunk_offsets.resize(nchunks + 1);
    PsCheck(ps_memcpy_d2h_async(chunk_offsets.data(), d_off, (nchunks + 1) * 8, s), "chunk offsets download");
    PsCheck(ps_stream_sync(s), "stream synchronize");
    bytes.resize(chunk_offsets.back());
    PsCheck(ps_memcpy_d2h_async(bytes.data(), d_out, bytes.size(), s), "stream download");
    PsCheck(ps_stream_sync(s), "stream synchronize");
}

}  // namespace fcstream

using namespace fcstream;

namespace {

// one chunk of `input` (chunk c of `chunks`) decoded into a host buffer
void DecodeChunk(const std::vector<uint8_t>& input, const std::vector<uint64_t>& chunks, size_t c, Buffer<uint32_t>& out,
                 DeviceBytes& dout, DeviceBytes& scratch) {
    std::vector<uint8_t> one(input.begin() + (long)chunks[c], input.begin() + (long)chunks[c + 1]);
    const uint64_t n = DecodeToDevice(one, dout, scratch, nullptr);
    ensure(n <= out.Capacity());
    PsCheck(ps_memcpy_d2h_async(out.Buf(), dout.Ptr(), n * 4, nullptr), "ids download");
    PsCheck(ps_stream_sync(nullptr), "stream synchronize");
    out.SetSize(n);
}

// reads the next Store chunk of the segment into input and finds its
// FrontierCompression chunks; false at the end of the segment
bool ReadStoreChunk(Store& store, int segment, std::vector<uint8_t>& input, std::vector<uint64_t>& chunks) {
    input.resize(XSEG_PIECE);
    const size_t n = store.Read(segment, input.data(), input.size());
    input.resize(n);
    if (n == 0) return false;
    ChunkBoundaries(input.data(), input.size(), chunks);
    return true;
}

}  // namespace

CompressedSegmentReader::CompressedSegmentReader(Store& store) : m_Store(store), m_IndexBuffer(MAX_CHUNK_VALUES) {}

void CompressedSegmentReader::SetSegment(int segment) {
    m_Segment = segment;
    m_Store.Rewind(segment);
    m_Input.clear();
    m_Chunks.assign(1, 0);
    m_Next = 0;
}

void CompressedSegmentReader::Delete(int segment) { m_Store.Delete(segment); }

bool CompressedSegmentReader::NextStoreChunk() {
    m_Next = 0;
    return ReadStoreChunk(m_Store, m_Segment, m_Input, m_Chunks);
}

// Common/CompressedFrontier.cpp:90-100
Buffer<uint32_t>& CompressedSegmentReader::Read() {
    m_IndexBuffer.Clear();
    if (m_Next + 1 >= m_Chunks.size() && !NextStoreChunk()) return m_IndexBuffer;
    DecodeChunk(m_Input, m_Chunks, m_Next++, m_IndexBuffer, m_DevOut, m_Scratch);
    return m_IndexBuffer;
}

uint64_t CompressedSegmentReader::ReadAllToDevice(DeviceBytes& ids, ps_stream s) {
    std::vector<uint8_t> all(m_Input.begin() + (long)m_Chunks[m_Next], m_Input.end());
    std::vector<uint8_t> piece;
    std::vector<uint64_t> chunks;
    while (ReadStoreChunk(m_Store, m_Segment, piece, chunks)) all.insert(all.end(), piece.begin(), piece.end());
    m_Input.clear();
    m_Chunks.assign(1, 0);
    m_Next = 0;
    return DecodeToDevice(all, ids, m_Scratch, s);
}

CompressedSegmentWriter::CompressedSegmentWriter(Store& store) : m_Store(store), m_IndexBuffer(MAX_CHUNK_VALUES) {}

void CompressedSegmentWriter::SetSegment(int segment) {
    m_Segment = segment;
    ensure(m_IndexBuffer.IsEmpty());
    ensure(m_Output.empty());
}

void CompressedSegmentWriter::Add(uint32_t value) {
    m_IndexBuffer.Add(value);
    if (m_IndexBuffer.IsFull()) FlushData();
}

// FrontierCompression::Encode of the buffered ids, appended to the output
// piece, which goes to the Store first if the chunk would not fit
// (Common/CompressedFrontier.cpp:142-153)
void CompressedSegmentWriter::FlushData() {
    const uint64_t n = m_IndexBuffer.Size();
    uint32_t* d = (uint32_t*)m_DevIds.Reserve(n * 4 + 64);
    PsCheck(ps_memcpy_h2d_async(d, m_IndexBuffer.Buf(), n * 4, nullptr), "ids upload");
    std::vector<uint8_t> bytes;
    std::vector<uint64_t> off;
    EncodeFromDevice(d, n, MAX_CHUNK_VALUES, false, bytes, off, m_Scratch, nullptr);
    m_IndexBuffer.Clear();
    if (m_Output.size() + bytes.size() > FRONTIER_PIECE) {
        m_Store.Write(m_Segment, m_Output.data(), m_Output.size());
        m_Output.clear();
    }
    m_Output.insert(m_Output.end(), bytes.begin(), bytes.end());
}

void CompressedSegmentWriter::Flush() {
    if (!m_IndexBuffer.IsEmpty()) FlushData();
    if (!m_Output.empty()) {
        m_Store.Write(m_Segment, m_Output.data(), m_Output.size());
        m_Output.clear();
    }
}

void CompressedSegmentWriter::AddFromDevice(const uint32_t* d_ids, uint64_t n, ps_stream s) {
    ensure(m_IndexBuffer.IsEmpty());  // the device ids start a new chunk
    std::vector<uint8_t> bytes;
    std::vector<uint64_t> off;
    EncodeFromDevice(d_ids, n, MAX_CHUNK_VALUES, false, bytes, off, m_Scratch, s);
    for (size_t c = 0; c + 1 < off.size(); c++) {
        const size_t len = off[c + 1] - off[c];
        if (m_Output.size() + len > FRONTIER_PIECE) {
            m_Store.Write(m_Segment, m_Output.data(), m_Output.size());
            m_Output.clear();
        }
        m_Output.insert(m_Output.end(), bytes.begin() + (long)off[c], bytes.begin() + (long)off[c + 1]);
    }
}

CompressedCrossSegmentReader::CompressedCrossSegmentReader(StoreSet& storeSet)
    : m_StoreSet(storeSet), m_IndexBuffer(MAX_CHUNK_VALUES) {}

void CompressedCrossSegmentReader::SetSegment(int segment) {
    m_Segment = segment;
    m_StoreSet.Rewind(segment);
    m_Input.clear();
    m_Chunks.assign(1, 0);
    m_Next = 0;
    m_LastOp = -1;
}

// Common/CompressedFrontier.cpp:155-166
Buffer<uint32_t>& CompressedCrossSegmentReader::Read(int op) {
    m_IndexBuffer.Clear();
    if (m_Next + 1 >= m_Chunks.size()) {
        m_Next = 0;
        if (!ReadStoreChunk(m_StoreSet.Stores[op], m_Segment, m_Input, m_Chunks)) return m_IndexBuffer;
        m_LastOp = op;
    }
    ensure(m_LastOp == op);
    DecodeChunk(m_Input, m_Chunks, m_Next++, m_IndexBuffer, m_DevOut, m_Scratch);
    return m_IndexBuffer;
}

void CompressedCrossSegmentReader::Delete(int segment) { m_StoreSet.Delete(segment); }
