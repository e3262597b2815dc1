// ose_entry.h -- registry of the compiled FFT extents (one FftEntry per N).
#pragma once
#include <cstddef>

namespace ose {

// one compiled launch shape of a strided (y/z) pass: T x columns per
// workgroup, split = 1 for the half-LDS (real, then imaginary) exchange
struct StridedKern {
    int T = 0, split = 0, threads = 0;
    size_t lds = 0;
    const void* fn = nullptr;
};

// one compiled launch shape of the x pass: P lines per workgroup, a kernel
// per XMode, split = 1 for the half-LDS exchange, twb = 1 for the
// scheduling barrier before the twiddle multiplies
struct XKern {
    int P = 0, split = 0, twb = 0, threads = 0;
    size_t lds = 0;
    const void* fn[9] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
};

constexpr int kMaxVariants = 12;

struct FftEntry {
    int N = 0, N1 = 0, N2 = 0;
    int ns = 0;                       // strided variants
    StridedKern yf[kMaxVariants];     // forward strided transform
    StridedKern yi[kMaxVariants];     // inverse strided transform
    StridedKern zs[kMaxVariants];     // forward * |k|/N * inverse
    int nx = 0;                       // x-pass variants
    XKern xk[kMaxVariants];
};

// nullptr if N is not a compiled extent
const FftEntry* find_entry(int N);
// smallest compiled extent >= n (0 if none)
int next_supported(int n);

}  // namespace ose
