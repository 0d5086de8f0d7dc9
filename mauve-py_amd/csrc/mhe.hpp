// mhe.hpp -- MatchHashEntry of k genomes and the comparisons AddHashEntry
// makes with it (LM/MatchHashEntry.{h,cpp}), for host code and device
// kernels alike.
#pragma once
#include <cstdint>
#include <cstddef>
#include "extend.hpp"   // MH_HD

namespace mh {

// A MatchHashEntry of up to MK genomes (start 0: absent), first = the index
// of its first present genome (FirstStart); starts of type T.
template <int MK, class T> struct EntT {
    T s[MK];
    uint64_t len;
    int64_t off;
    uint64_t mersize;
    int first;
};
using EntK = EntT<4, int64_t>;     // PairwiseMatchFinder, up to four genomes
using EntU = EntT<16, int32_t>;    // UniqueMatchFinder, up to sixteen

// The loops below run over all MK genome slots with the genome range as a
// condition (unrolled), never from a run-time start index: on the GPU an
// entry indexed at run time lives in scratch memory.

// MatchHashEntry::Contains (a contains b), LM/MatchHashEntry.cpp:164-200
template <class E> MH_HD inline bool contains_k(const E &a, const E &b, int k)
{
    constexpr int MK = (int)(sizeof(a.s) / sizeof(a.s[0]));
    if (a.off != b.off) return false;
    const int f = b.first;
    int64_t bf = 0, af = 0;
    MH_PRAGMA_UNROLL
    for (int g = 0; g < MK; ++g)
        if (g == f) {
            bf = b.s[g];
            af = a.s[g];
        }
    const int64_t diff = bf - af;
    if (af == 0) return false;
    if (diff < 0 || a.len < b.len + (uint64_t)diff) return false;
    const int64_t diff_rc = (int64_t)b.len - (int64_t)a.len + diff;
    bool ok = true;
    MH_PRAGMA_UNROLL
    for (int g = 0; g < MK; ++g) {
        if (g <= f || g >= k) continue;
        const int64_t di = (int64_t)b.s[g] - (int64_t)a.s[g];
        if (b.s[g] == 0 && a.s[g] == 0) continue;
        if (b.s[g] < 0 && diff_rc == di) continue;
        if (diff != di) ok = false;
    }
    return ok;
}

// strict_start_lessthan_ptr, LM/MatchHashEntry.cpp:48-67
template <class E> MH_HD inline bool strict_less_k(const E &a, const E &b, int k)
{
    constexpr int MK = (int)(sizeof(a.s) / sizeof(a.s[0]));
    if (a.first == b.first) {
        int res = 0;   // 1: a first, 2: b first
        MH_PRAGMA_UNROLL
        for (int g = 0; g < MK; ++g) {
            if (g >= k || res) continue;
            int64_t as = a.s[g], bs = b.s[g];
            if (as < 0) as = (int64_t)((uint64_t)(-as) + a.len - a.mersize);
            if (bs < 0) bs = (int64_t)((uint64_t)(-bs) + b.len - b.mersize);
            if (as != bs) res = as < bs ? 1 : 2;
        }
        if (res) return res == 1;
    }
    return a.first < b.first;
}

// MheCompare, LM/MatchHashEntry.h:121-143
template <class E> MH_HD inline bool mhe_less_k(const E &a, const E &b, int k)
{
    constexpr int MK = (int)(sizeof(a.s) / sizeof(a.s[0]));
    if (a.first > b.first) return true;
    if (a.first != b.first) return false;
    int res = 0;   // the first genome from a.first on present in one only: 1 in b, 2 in a
    MH_PRAGMA_UNROLL
    for (int g = 0; g < MK; ++g) {
        if (g < a.first || g >= k || res) continue;
        if (a.s[g] == 0 && b.s[g] != 0) res = 1;
        else if (a.s[g] != 0 && b.s[g] == 0) res = 2;
    }
    if (res) return res == 1;
    if (contains_k(a, b, k) || contains_k(b, a, k)) return false;
    return strict_less_k(a, b, k);
}

// std::lower_bound as libstdc++ implements it (bits/stl_algobase.h): the
// probe sequence matters because MheCompare is not a strict weak order.
// at(i) gives element i of the sorted range [0, n).
template <class E, class At> MH_HD inline size_t lower_bound_at(size_t n, const E &q, int k, At at)
{
    size_t first = 0, len = n;
    while (len > 0) {
        const size_t half = len >> 1, mid = first + half;
        if (mhe_less_k(at(mid), q, k)) {
            first = mid + 1;
            len = len - half - 1;
        } else {
            len = half;
        }
    }
    return first;
}

// CalculateOffset, LM/MatchHashEntry.cpp:141-160 (Length is the seed length
// when HashMatch calls it)
template <class E> MH_HD inline int64_t offset_k(const E &e, int k)
{
    int64_t off = 0;
    const int64_t ref = e.s[e.first];
    for (int i = e.first + 1; i < k; ++i)
        if (e.s[i] != 0) off += (int64_t)e.s[i] - ref - (e.s[i] < 0 ? (int64_t)e.len : 0);
    return off;
}

}  // namespace mh
