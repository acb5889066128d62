// Segment stores of the disk-based drivers, with the interface of the
// reference's Common/Store.h:13-176: a Store keeps, per segment, a sequence of
// byte chunks; Write appends a chunk, Read returns the next chunk of the
// segment (whole, into a buffer at least as large), Rewind restarts it, Delete
// drops it.  Chunks live in files (POSIX), split over directories and files as
// StoreOptions says; the chunk index stays in memory.
#pragma once

#include <cstddef>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "Buffer.h"
#include "Util.h"

// Common/Store.h:18-29
struct StoreOptions {
    // segments are spread over these directories round-robin
    std::vector<std::string> directories = {"."};
    // 0: a file per segment; 1: one file per directory; n > 1: n files per
    // directory, each holding a consecutive range of segments
    int filesPerPath = 5;
};

class StoreImpl {
public:
    virtual ~StoreImpl() noexcept {}
    virtual int MaxSegments() const = 0;
    virtual uint64_t Length(int segment) const = 0;
    virtual bool HasData(int segment) const = 0;
    virtual void Rewind(int segment) = 0;
    virtual void Write(int segment, const void* buffer, size_t size) = 0;
    virtual size_t Read(int segment, void* buffer, size_t size) = 0;
    virtual void Delete(int segment) = 0;
    virtual uint64_t TotalLength() const {
        uint64_t t = 0;
        for (int i = 0; i < MaxSegments(); i++) t += Length(i);
        return t;
    }
    virtual void RewindAll() {
        for (int i = 0; i < MaxSegments(); i++) Rewind(i);
    }
    virtual void DeleteAll() {
        for (int i = 0; i < MaxSegments(); i++) Delete(i);
    }
};

class Store {
public:
    static Store CreateSingleFileStore(int maxSegments, const std::string& filePath);
    static Store CreateMultiFileStore(int maxSegments, const std::string& directory, const std::string& prefix);
    static Store CreateSequentialStore(int maxSegments, std::vector<std::string> filePaths);
    static Store CreateFileStore(int maxSegments, const std::string& name, StoreOptions options);

    Store(Store&&) noexcept = default;
    Store& operator=(Store&&) noexcept = default;

    int MaxSegments() const { return m_Impl->MaxSegments(); }
    bool HasData(int segment) const { return m_Impl->HasData(segment); }
    uint64_t Length(int segment) const { return m_Impl->Length(segment); }
    uint64_t TotalLength() const { return m_Impl->TotalLength(); }
    void Rewind(int segment) { m_Impl->Rewind(segment); }
    void RewindAll() { m_Impl->RewindAll(); }
    // appends one chunk (nothing for size 0)
    void Write(int segment, const void* buffer, size_t size);
    // the next chunk of the segment, or 0 bytes at its end; the buffer must
    // hold the chunk
    size_t Read(int segment, void* buffer, size_t size);
    void Delete(int segment) { m_Impl->Delete(segment); }
    void DeleteAll() { m_Impl->DeleteAll(); }

    template <typename T>
    void WriteArray(int segment, const T* buffer, size_t size) {
        Write(segment, (const void*)buffer, size * sizeof(T));
    }
    template <typename T>
    void Write(int segment, const std::vector<T>& buffer) {
        WriteArray<T>(segment, buffer.data(), buffer.size());
    }
    template <typename T>
    void Write(int segment, const Buffer<T>& buffer) {
        WriteArray<T>(segment, buffer.Buf(), buffer.Size());
    }
    template <typename T>
    size_t ReadArray(int segment, T* buffer, size_t capacity) {
        const size_t read = Read(segment, (void*)buffer, capacity * sizeof(T));
        ensure(read % sizeof(T) == 0);
        return read / sizeof(T);
    }
    template <typename T>
    void Read(int segment, std::vector<T>& buffer) {
        buffer.resize(buffer.capacity());
        buffer.resize(ReadArray<T>(segment, buffer.data(), buffer.capacity()));
    }
    template <typename T>
    void Read(int segment, Buffer<T>& buffer) {
        buffer.SetSize(ReadArray<T>(segment, buffer.Buf(), buffer.Capacity()));
    }

private:
    explicit Store(std::unique_ptr<StoreImpl> impl) : m_Impl(std::move(impl)) {}
    std::unique_ptr<StoreImpl> m_Impl;
};

// Common/Store.h:140-176: one store per operator (the cross-segment buckets)
struct StoreSet {
    std::vector<Store> Stores;

    int MaxSegments() const { return Stores[0].MaxSegments(); }
    uint64_t TotalLength() const {
        uint64_t t = 0;
        for (const auto& s : Stores) t += s.TotalLength();
        return t;
    }
    void Rewind(int segment) {
        for (auto& s : Stores) s.Rewind(segment);
    }
    void RewindAll() {
        for (auto& s : Stores) s.RewindAll();
    }
    void Delete(int segment) {
        for (auto& s : Stores) s.Delete(segment);
    }
    void DeleteAll() {
        for (auto& s : Stores) s.DeleteAll();
    }
    void Swap(StoreSet& other) {
        for (size_t i = 0; i < Stores.size(); i++) std::swap(Stores[i], other.Stores[i]);
    }
};
