// HanoiSearch<size> with the reference's interface (HanoiTowers/HanoiSearch.h),
// run by the device-resident vertical search (ps_hanoi_search_*): the segments
// a layer reaches are held in HBM instead of disk stores, so `threads` and
// `directories` are accepted and not used.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

struct SearchOptions {
    int maxSteps = 1000000;
    int threads = 4;
    std::vector<std::string> directories = {"."};
};

// three-bit BFS of the reduced states from initialState; result[l] is layer
// l's count restored to the unreduced puzzle (result[0] = 1), for at most
// maxSteps layers after the first (HanoiTowers/HanoiSearch.cpp:215-290)
template <int size>
std::vector<uint64_t> HanoiSearch(std::string initialState, SearchOptions options);
