// Fixed-size zero-initialised host array with the interface of the reference's
// TArray<T> (Common/Array.h:9-64): Size, SetSize (reallocates and zeroes),
// operator[], Buf, Clear; movable, not copyable.
#pragma once

#include <cstddef>
#include <cstring>
#include <memory>
#include <utility>

#include "Util.h"

template <typename T>
class TArray {
public:
    TArray() = default;
    explicit TArray(size_t size) { SetSize(size); }

    TArray(const TArray&) = delete;
    TArray& operator=(const TArray&) = delete;
    TArray(TArray&& other) noexcept { swap(other); }
    TArray& operator=(TArray&& other) noexcept {
        swap(other);
        return *this;
    }

    size_t Size() const { return m_Size; }

    // Drops the contents: the new array is all zero bytes.
    void SetSize(size_t size) {
        m_Data.reset();
        m_Size = 0;
        if (size) {
            m_Data.reset(new T[size]);
            std::memset((void*)m_Data.get(), 0, size * sizeof(T));
        }
        m_Size = size;
    }

    T& operator[](size_t i) noexcept { return m_Data[i]; }
    const T& operator[](size_t i) const noexcept { return m_Data[i]; }

    T* Buf() { return m_Data.get(); }
    const T* Buf() const { return m_Data.get(); }

    void Clear() {
        if (m_Size) std::memset((void*)m_Data.get(), 0, m_Size * sizeof(T));
    }

private:
    void swap(TArray& o) noexcept {
        std::swap(m_Size, o.m_Size);
        std::swap(m_Data, o.m_Data);
    }

    size_t m_Size = 0;
    std::unique_ptr<T[]> m_Data;
};
