/*
 * or_codecs.c -- CPU restatement of the reference frontier codecs: StreamVInt,
 * OpBitsCompression<2/4/8> and FrontierCompression (the oracle; test
 * infrastructure only, see oracle.h).  Byte layouts are reproduced exactly.
 */
#include "oracle.h"

#include <string.h>

static inline uint32_t rd32(const uint8_t* p) {
    uint32_t v;
    memcpy(&v, p, 4);
    return v;
}
static inline void wr32(uint8_t* p, uint32_t v) { memcpy(p, &v, 4); }
static inline uint64_t rd64(const uint8_t* p) {
    uint64_t v;
    memcpy(&v, p, 8);
    return v;
}
static inline void wr64(uint8_t* p, uint64_t v) { memcpy(p, &v, 8); }

/* ------------------------------------------------------------ StreamVInt */

/* Common/StreamVInt.cpp:94-141,185-196: one descriptor byte (value j's byte
 * length minus one at bits 2j..2j+1), then each value's low bytes, little
 * endian.  The length of a value is the position of its highest nonzero byte
 * (at least one byte). */
int or_svint_encode_tuple(const uint32_t* v4, uint8_t* buf) {
    uint8_t desc = 0;
    int pos = 1;
    for (int j = 0; j < 4; j++) {
        uint32_t v = v4[j];
        int len = (v >> 24) ? 4 : (v >> 16) ? 3 : (v >> 8) ? 2 : 1;
        desc |= (uint8_t)((len - 1) << (2 * j));
        for (int b = 0; b < len; b++) buf[pos++] = (uint8_t)(v >> (8 * b));
    }
    buf[0] = desc;
    return pos;
}

/* Common/StreamVInt.cpp:143-183,198-206 */
int or_svint_decode_tuple(const uint8_t* buf, uint32_t* v4) {
    uint8_t desc = buf[0];
    int pos = 1;
    for (int j = 0; j < 4; j++) {
        int len = 1 + ((desc >> (2 * j)) & 3);
        uint32_t v = 0;
        for (int b = 0; b < len; b++) v |= (uint32_t)buf[pos++] << (8 * b);
        v4[j] = v;
    }
    return pos;
}

/* Common/StreamVInt.cpp:208-216: eight op-flag bytes folded into four. */
void or_svint_encode_bounds_tuple(const uint8_t* b8, uint8_t* buf) {
    uint64_t all = rd64(b8);
    wr32(buf, (uint32_t)(all | (all >> 28)));
}

void or_svint_decode_bounds_tuple(const uint8_t* buf, uint8_t* b8) {
    uint64_t stored = rd32(buf);
    wr64(b8, (stored & 0x0F0F0F0Full) | ((stored & 0xF0F0F0F0ull) << 28));
}

/* Common/StreamVInt.cpp:234-262 */
int or_svint_encode_indexes(int count, const uint32_t* idx, uint8_t* buf) {
    int pos = 0;
    for (int i = 0; i < count; i += 4) pos += or_svint_encode_tuple(&idx[i], &buf[pos]);
    return pos;
}

int or_svint_decode_indexes(int count, const uint8_t* buf, uint32_t* idx) {
    int pos = 0;
    for (int i = 0; i < count; i += 4) pos += or_svint_decode_tuple(&buf[pos], &idx[i]);
    return pos;
}

int or_svint_decode_indexes_diff(int count, const uint8_t* buf, uint32_t* idx) {
    int pos = 0;
    uint32_t cur = 0;
    for (int i = 0; i < count; i += 4) {
        pos += or_svint_decode_tuple(&buf[pos], &idx[i]);
        for (int j = 0; j < 4; j++) cur = (idx[i + j] += cur);
    }
    return pos;
}

/* Common/StreamVInt.cpp:278-300: u32 count, then the delta-coded values padded
 * with zeros to a multiple of 8.  The input array is turned into deltas. */
int or_svint_encode(int count, uint32_t* idx, uint8_t* buf) {
    if (count == 0) return 0;
    int adjusted = (count + 7) / 8 * 8;
    for (int i = count - 1; i > 0; i--) idx[i] -= idx[i - 1];
    for (int i = count; i < adjusted; i++) idx[i] = 0;
    wr32(buf, (uint32_t)count);
    return 4 + or_svint_encode_indexes(adjusted, idx, buf + 4);
}

/* Common/StreamVInt.cpp:233-251: *size receives the bytes consumed. */
int or_svint_decode(int* size, const uint8_t* buf, uint32_t* idx) {
    if (*size == 0) return 0;
    int count = (int)rd32(buf);
    int adjusted = (count + 7) / 8 * 8;
    *size = 4 + or_svint_decode_indexes_diff(adjusted, buf + 4, idx);
    return count;
}

/* Common/StreamVInt.cpp:253-299 */
int or_svint_encode_with_bounds(int count, uint32_t* idx, uint8_t* bounds, uint8_t* buf) {
    if (count == 0) return 0;
    int adjusted = (count + 7) / 8 * 8;
    for (int i = count - 1; i > 0; i--) idx[i] -= idx[i - 1];
    for (int i = count; i < adjusted; i++) {
        idx[i] = 0;
        bounds[i] = 0;
    }
    wr32(buf, (uint32_t)count);
    int pos = 4 + or_svint_encode_indexes(adjusted, idx, buf + 4);
    for (int i = 0; i < adjusted; i += 8) or_svint_encode_bounds_tuple(&bounds[i], &buf[pos + i / 2]);
    return pos + adjusted / 2;
}

int or_svint_decode_with_bounds(int* size, const uint8_t* buf, uint32_t* idx, uint8_t* bounds) {
    if (*size == 0) return 0;
    int count = (int)rd32(buf);
    int adjusted = (count + 7) / 8 * 8;
    int pos = 4 + or_svint_decode_indexes_diff(adjusted, buf + 4, idx);
    for (int i = 0; i < adjusted; i += 8) or_svint_decode_bounds_tuple(&buf[pos + i / 2], &bounds[i]);
    *size = pos + adjusted / 2;
    return count;
}

/* ------------------------------------------------------------ OpBits */

/* Common/OpBitsCompression.cpp:10-76: eight op-flag bytes per tuple folded into
 * BITS bytes.  count must be a multiple of 8 (the reference stops otherwise). */
int64_t or_opbits_encode(int bits, uint64_t count, const uint8_t* ops, uint8_t* buf) {
    if (count & 7) return -1;
    uint64_t pos = 0;
    for (uint64_t i = 0; i < count; i += 8) {
        uint64_t all = rd64(&ops[i]);
        if (bits == 2) {
            uint16_t v = (uint16_t)(all | (all >> 14) | (all >> 28) | (all >> 42));
            memcpy(&buf[pos], &v, 2);
            pos += 2;
        } else if (bits == 4) {
            wr32(&buf[pos], (uint32_t)(all | (all >> 28)));
            pos += 4;
        } else if (bits == 8) {
            wr64(&buf[pos], all);
            pos += 8;
        } else {
            return -1;
        }
    }
    return (int64_t)pos;
}

int64_t or_opbits_decode(int bits, uint64_t count, const uint8_t* buf, uint8_t* ops) {
    if (count & 7) return -1;
    uint64_t pos = 0;
    for (uint64_t i = 0; i < count; i += 8) {
        uint64_t out;
        if (bits == 2) {
            uint16_t s16;
            memcpy(&s16, &buf[pos], 2);
            uint64_t s = s16;
            out = (s & 0x0303ull) | ((s & 0x0C0Cull) << 14) | ((s & 0x3030ull) << 28) | ((s & 0xC0C0ull) << 42);
            pos += 2;
        } else if (bits == 4) {
            uint64_t s = rd32(&buf[pos]);
            out = (s & 0x0F0F0F0Full) | ((s & 0xF0F0F0F0ull) << 28);
            pos += 4;
        } else if (bits == 8) {
            out = rd64(&buf[pos]);
            pos += 8;
        } else {
            return -1;
        }
        wr64(&ops[i], out);
    }
    return (int64_t)pos;
}

/* ------------------------------------------------- FrontierCompression */

#define CODEC_MAP 0x80000000u

int or_fc_is_bitmap(const uint8_t* buf) { return (rd32(buf) & CODEC_MAP) != 0; }

/* Common/FrontierCompression.cpp:28-34 (SIZE_MAX when unsorted at the ends) */
int64_t or_fc_bitmap_size(int count, const uint32_t* idx) {
    uint64_t first = idx[0], last = idx[count - 1];
    if (last < first) return -1;
    return (int64_t)((last - first + 63) / 64 * 8 + 8);
}

/* Common/FrontierCompression.cpp:36-51: header (words | 0x80000000), first,
 * then a bitmap of value - first - 1.  Returns -1 where the reference would
 * write out of bounds (unsorted input; it never calls this then). */
int or_fc_encode_bitmap(int count, const uint32_t* idx, uint8_t* buf) {
    uint64_t first = idx[0], last = idx[count - 1];
    if (last < first) return -1;
    uint64_t words = (last - first + 63) / 64;
    wr32(buf, (uint32_t)(words | CODEC_MAP));
    wr32(buf + 4, (uint32_t)first);
    memset(buf + 8, 0, words * 8);
    for (int i = 1; i < count; i++) {
        uint64_t val = (uint64_t)idx[i] - first - 1;
        if (val >= words * 64) return -1;
        buf[8 + val / 8] |= (uint8_t)(1u << (val % 8));
    }
    return (int)(8 + words * 8);
}

/* Common/FrontierCompression.cpp:53-72 */
int or_fc_encode_bitmap_with_check(int count, const uint32_t* idx, uint8_t* buf) {
    uint64_t first = idx[0], last = idx[count - 1];
    if (last <= first) return -1;
    uint64_t max = last - first - 1;
    uint64_t words = (last - first + 63) / 64;
    wr32(buf, (uint32_t)(words | CODEC_MAP));
    wr32(buf + 4, (uint32_t)first);
    memset(buf + 8, 0, words * 8);
    for (int i = 1; i < count; i++) {
        uint64_t val = (uint64_t)idx[i] - first - 1;
        if (val > max) return -1;
        buf[8 + val / 8] |= (uint8_t)(1u << (val % 8));
    }
    return (int)(8 + words * 8);
}

/* Common/FrontierCompression.cpp:74-90 */
int or_fc_decode_bitmap(int* size, const uint8_t* buf, uint32_t* idx) {
    uint32_t words = rd32(buf) & ~CODEC_MAP;
    uint32_t first = rd32(buf + 4);
    int n = 0;
    idx[n++] = first;
    for (uint64_t w = 0; w < words; w++) {
        uint64_t v = rd64(buf + 8 + w * 8);
        while (v) {
            idx[n++] = first + (uint32_t)(w * 64 + (uint64_t)__builtin_ctzll(v)) + 1;
            v &= v - 1;
        }
    }
    *size = (int)(8 + words * 8);
    return n;
}

/* size_t arithmetic of Common/FrontierCompression.cpp:92-102 (BitMapSize()*8
 * wraps for unsorted input and then never selects the bitmap). */
static uint64_t bitmap_bits(int count, const uint32_t* idx) {
    int64_t s = or_fc_bitmap_size(count, idx);
    uint64_t bytes = s < 0 ? UINT64_MAX : (uint64_t)s;
    return bytes * 8;
}

int or_fc_encode(int count, uint32_t* idx, uint8_t* buf) {
    if (count == 0) return 0;
    uint64_t min_bits = (uint64_t)count * 10;
    if (count > 128 && bitmap_bits(count, idx) < min_bits - 8) return or_fc_encode_bitmap(count, idx, buf);
    return or_svint_encode(count, idx, buf);
}

/* Common/FrontierCompression.cpp:104-115 */
int or_fc_encode_with_check(int count, uint32_t* idx, uint8_t* buf) {
    if (count == 0) return 0;
    uint64_t min_bits = (uint64_t)count * 10;
    if (count > 128 && bitmap_bits(count, idx) < min_bits) {
        int r = or_fc_encode_bitmap_with_check(count, idx, buf);
        if (r >= 0) return r;
    }
    return or_svint_encode(count, idx, buf);
}

/* Common/FrontierCompression.cpp:117-125 */
int or_fc_decode(int* size, const uint8_t* buf, uint32_t* idx) {
    if (*size == 0) return 0;
    if (or_fc_is_bitmap(buf)) return or_fc_decode_bitmap(size, buf, idx);
    return or_svint_decode(size, buf, idx);
}
