// gsap/radixsort.h -- drop-in for nvRadixSort::RadixSort
// (physicboxSap2.16_Fermi/sapDetection/radixsort.h:110-278), header-only over
// gsap_radix_sort_* in gsap.h.  Stable LSD sort in place on device arrays.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdio>

#include "../gsap.h"

namespace nvRadixSort {

class RadixSort {
public:
    RadixSort(unsigned int maxElements, bool keysOnly = false) {
        if (gsap_radix_sort_create(maxElements, keysOnly ? 1 : 0, &m_s) != GSAP_OK)
            std::fprintf(stderr, "RadixSort: %s\n", gsap_last_error());
    }
    ~RadixSort() { gsap_radix_sort_destroy(m_s); }
    RadixSort(const RadixSort&) = delete;
    RadixSort& operator=(const RadixSort&) = delete;

    // sort(uint* keys, uint* values, uint n, uint keyBits) (radixsort.h:151-178)
    void sort(unsigned int* keys, unsigned int* values, unsigned int numElements,
              unsigned int keyBits) {
        report(gsap_radix_sort_u32(m_s, m_stream, keys, values, numElements, keyBits));
    }
    // sort(float* keys, uint* values, uint n, uint keyBits, bool negativeKeys)
    // (radixsort.h:180-198)
    void sort(float* keys, unsigned int* values, unsigned int numElements, unsigned int keyBits,
              bool negativeKeys) {
        report(gsap_radix_sort_f32(m_s, m_stream, keys, values, numElements, keyBits,
                                   negativeKeys ? 1 : 0));
    }
    void setStream(hipStream_t s) { m_stream = s; }

private:
    static void report(int rc) {
        if (rc != GSAP_OK) std::fprintf(stderr, "RadixSort::sort: %s\n", gsap_last_error());
    }
    gsap_radix_sort* m_s = nullptr;
    hipStream_t m_stream = nullptr;
};

}  // namespace nvRadixSort
