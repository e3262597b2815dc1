// ose_entry.h -- registry of the compiled FFT extents (one FftEntry per N).
#pragma once
#include <cstddef>

namespace ose {

struct FftEntry {
    int N = 0, N1 = 0, N2 = 0;
    // strided (y/z) passes: T columns per workgroup
    int s_T = 0, s_threads = 0;
    size_t s_lds = 0;
    const void* ystr[2] = {nullptr, nullptr};   // [0] forward, [1] inverse
    const void* zsym = nullptr;                 // forward-z * |k|/N * inverse-z
    // x-line pass: P lines per workgroup, one kernel per XMode
    int x_P = 0, x_threads = 0;
    size_t x_lds = 0;
    const void* xk[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
};

// nullptr if N is not a compiled extent
const FftEntry* find_entry(int N);
// smallest compiled extent >= n (0 if none)
int next_supported(int n);

}  // namespace ose
