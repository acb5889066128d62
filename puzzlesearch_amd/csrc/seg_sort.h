// Device prefix sums and stable counting sorts shared by the segment steps
// and the bucketing codec entry point.  Hand-written for wave64: every sort is
// a sequence of stable counting passes (per-block key histograms, one prefix
// sum over the key-major histogram matrix, then a scatter in which one wave
// walks its block's entries in order, ranks equal keys with ballots, and keeps
// a cursor per key in LDS), so entries with equal keys keep their input order.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace ps {

// Exclusive prefix sums of n values: out[i] = in[0] + ... + in[i-1], and
// out[n] = the total (n + 1 entries).  In stream order; no wait.
int scan_excl_u32(const uint32_t* d_in, uint64_t n, unsigned long long* d_out, hipStream_t s);
int scan_excl_u64(const unsigned long long* d_in, uint64_t n, unsigned long long* d_out, hipStream_t s);

// Groups m entries kv[i] = key << 32 | value by key (key < key_space) and
// keeps their input order inside every key: d_values_out receives the values
// key by key in ascending key order, `keys` the nonempty keys ascending and
// `counts` their sizes.  Entries whose key is >= key_space are left out and
// counted in *dropped.  m < 2^31.  Waits for the stream (the key list is
// returned to the host).
int stable_group_by_key(const unsigned long long* d_kv, uint64_t m, uint32_t key_space, uint32_t* d_values_out,
                        std::vector<uint32_t>& keys, std::vector<uint64_t>& counts, uint64_t* dropped,
                        hipStream_t s);

// Sorts m entries kv by value (the low `value_bits` bits, ascending), keeping
// the input order of equal values, in place of d_kv; d_tmp holds m entries.
// An 8-bit stable counting pass per byte of the value.  In stream order.
int stable_sort_by_value(unsigned long long* d_kv, unsigned long long* d_tmp, uint64_t m, int value_bits,
                         hipStream_t s);

}  // namespace ps
