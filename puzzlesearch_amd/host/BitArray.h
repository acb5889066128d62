// Host bit array with the interface of Common/BitArray.h:69-169 (bit i in
// 64-bit word i/64 at bit i%64), used by the reference-style host drivers.
// The device-resident drivers keep the same layout in HBM.
#pragma once

#include <cstdint>
#include <vector>

class BitArray {
public:
    BitArray() = default;
    explicit BitArray(uint64_t size) : m_Values((size + 63) / 64, 0) {}

    void Resize(uint64_t size) { m_Values.resize((size + 63) / 64); }
    void Set(uint64_t i) { m_Values[i >> 6] |= 1ull << (i & 63); }
    void Clear(uint64_t i) { m_Values[i >> 6] &= ~(1ull << (i & 63)); }
    bool Get(uint64_t i) const { return (m_Values[i >> 6] >> (i & 63)) & 1; }

    void AndNot(const BitArray& e) {
        for (size_t i = 0; i < m_Values.size(); i++) m_Values[i] &= ~e.m_Values[i];
    }
    void Or(const BitArray& o) {
        for (size_t i = 0; i < m_Values.size(); i++) m_Values[i] |= o.m_Values[i];
    }
    uint64_t BitsCount() const {
        uint64_t r = 0;
        for (uint64_t v : m_Values) r += (uint64_t)__builtin_popcountll(v);
        return r;
    }
    uint64_t AndNotAndCount(const BitArray& e) {
        uint64_t r = 0;
        for (size_t i = 0; i < m_Values.size(); i++) {
            m_Values[i] &= ~e.m_Values[i];
            r += (uint64_t)__builtin_popcountll(m_Values[i]);
        }
        return r;
    }
    void Clear() { std::fill(m_Values.begin(), m_Values.end(), 0ull); }

    size_t DataSize() const { return m_Values.size(); }
    uint64_t* Data() { return m_Values.data(); }
    const uint64_t* Data() const { return m_Values.data(); }

    template <typename F>
    void ScanBits(F func) const {
        for (size_t w = 0; w < m_Values.size(); w++) {
            uint64_t v = m_Values[w];
            while (v) {
                func((uint64_t)w * 64 + (uint64_t)__builtin_ctzll(v));
                v &= v - 1;
            }
        }
    }

    template <typename F>
    void ScanBitsAndClear(F func) {
        for (size_t w = 0; w < m_Values.size(); w++) {
            uint64_t v = m_Values[w];
            if (!v) continue;
            m_Values[w] = 0;
            while (v) {
                func((uint64_t)w * 64 + (uint64_t)__builtin_ctzll(v));
                v &= v - 1;
            }
        }
    }

private:
    std::vector<uint64_t> m_Values;
};
