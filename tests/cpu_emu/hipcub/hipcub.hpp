// CPU stand-in for the two hipCUB device algorithms the library calls (TEST
// INFRASTRUCTURE, see hip/hip_runtime.h in this directory): a stable LSD radix sort of
// key/value pairs and inclusive / exclusive prefix sums, with the same signatures and
// the same two-phase temporary-storage protocol.
#pragma once
#include <algorithm>
#include <cstdint>
#include <numeric>
#include <vector>

#include "../hip/hip_runtime.h"

namespace hipcub {

struct DeviceRadixSort {
    template <class K, class V>
    static hipError_t SortPairs(void* temp, size_t& bytes, const K* kin, K* kout, const V* vin, V* vout, int n,
                                int begin_bit = 0, int end_bit = (int)sizeof(K) * 8, hipStream_t = nullptr,
                                bool = false) {
        if (!temp) {
            bytes = 256;
            return hipSuccess;
        }
        if (bytes < 256) return hipErrorInvalidValue;
        std::vector<int> idx(n);
        std::iota(idx.begin(), idx.end(), 0);
        const unsigned long long mask =
            (end_bit - begin_bit >= 64) ? ~0ull : (((1ull << (end_bit - begin_bit)) - 1) << begin_bit);
        std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) {
            return ((unsigned long long)kin[a] & mask) < ((unsigned long long)kin[b] & mask);
        });
        std::vector<K> ko(n);
        std::vector<V> vo(n);
        for (int i = 0; i < n; i++) { ko[i] = kin[idx[i]]; vo[i] = vin[idx[i]]; }
        std::copy(ko.begin(), ko.end(), kout);
        std::copy(vo.begin(), vo.end(), vout);
        return hipSuccess;
    }
};

struct DeviceScan {
    template <class In, class Out>
    static hipError_t InclusiveSum(void* temp, size_t& bytes, In in, Out out, int n, hipStream_t = nullptr,
                                   bool = false) {
        if (!temp) { bytes = 256; return hipSuccess; }
        using T = typename std::remove_reference<decltype(out[0])>::type;
        T acc = T(0);
        for (int i = 0; i < n; i++) { acc += in[i]; out[i] = acc; }
        return hipSuccess;
    }
    template <class In, class Out>
    static hipError_t ExclusiveSum(void* temp, size_t& bytes, In in, Out out, int n, hipStream_t = nullptr,
                                   bool = false) {
        if (!temp) { bytes = 256; return hipSuccess; }
        using T = typename std::remove_reference<decltype(out[0])>::type;
        T acc = T(0);
        for (int i = 0; i < n; i++) { const T v = in[i]; out[i] = acc; acc += v; }
        return hipSuccess;
    }
};

}  // namespace hipcub
