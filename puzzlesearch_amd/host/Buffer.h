// Host buffer of constant capacity whose size changes without initialisation,
// with the interface of the reference's Buffer<T> (Common/Buffer.h:10-60): the
// type Puzzle::CrossSegmentPostProcess and the segment readers/writers pass
// cross-segment ids and compressed bytes in.
#pragma once

#include <cstddef>
#include <cstring>

#include "Array.h"
#include "Util.h"

template <typename T>
class Buffer {
public:
    explicit Buffer(size_t capacity) : m_Array(capacity) {}

    size_t Capacity() const { return m_Array.Size(); }
    // Reallocates (contents dropped, size kept only if it still fits).
    void SetCapacity(size_t capacity) {
        m_Array.SetSize(capacity);
        if (m_Size > capacity) m_Size = capacity;
    }

    size_t Size() const { return m_Size; }
    void SetSize(size_t size) {
        ensure(size <= Capacity());
        m_Size = size;
    }

    T& operator[](size_t index) { return m_Array[index]; }
    const T& operator[](size_t index) const { return m_Array[index]; }

    void Add(T value) {
        ensure(m_Size < Capacity());
        m_Array[m_Size++] = value;
    }
    void Clear() { m_Size = 0; }
    bool IsFull() const { return m_Size == Capacity(); }
    bool IsEmpty() const { return m_Size == 0; }

    TArray<T>& Array() { return m_Array; }
    const TArray<T>& Array() const { return m_Array; }
    T* Buf() { return m_Array.Buf(); }
    const T* Buf() const { return m_Array.Buf(); }

    void SetAllZero() { m_Array.Clear(); }

    bool CanAppend(const Buffer<T>& buf) const { return Size() + buf.Size() <= Capacity(); }
    void Append(const Buffer<T>& buf) {
        ensure(CanAppend(buf));
        if (buf.Size()) std::memcpy((void*)(Buf() + Size()), (const void*)buf.Buf(), buf.Size() * sizeof(T));
        m_Size += buf.Size();
    }

private:
    size_t m_Size = 0;
    TArray<T> m_Array;
};
