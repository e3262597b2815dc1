// ose_entry.h -- registry of the compiled FFT extents (one FftEntry per N).
#pragma once
#include <cstddef>

namespace ose {

struct GenPlan;

// one compiled launch shape of a strided (y/z) pass: T x columns per
// workgroup, split = 1 for the half-LDS (real, then imaginary) exchange,
// chunk > 0 for the chunked z pass (k_zsym_chunk: that many consecutive
// tiles per workgroup, each tile's successor prefetched; z symbol only, the
// slot's forward and inverse y kernels are empty)
struct StridedKern {
    int T = 0, split = 0, threads = 0, chunk = 0;
    size_t lds = 0;
    const void* fn = nullptr;
    const GenPlan* gp = nullptr;      // runtime-length kernels: their plan
};

// one compiled launch shape of the x pass: P lines per workgroup, a kernel
// per XMode, split = 1 for the half-LDS exchange, twb = 1 for the
// scheduling barrier before the twiddle multiplies
struct XKern {
    int P = 0, split = 0, twb = 0, threads = 0;
    size_t lds = 0;
    const void* fn[9] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    const GenPlan* gp = nullptr;      // runtime-length kernels: their plan
};

constexpr int kMaxVariants = 12;

struct FftEntry {
    int N = 0, N1 = 0, N2 = 0;        // N1 = N2 = 0: the runtime-length path
    int ns = 0;                       // strided variants
    StridedKern yf[kMaxVariants];     // forward strided transform
    StridedKern yi[kMaxVariants];     // inverse strided transform
    StridedKern zs[kMaxVariants];     // forward * |k|/N * inverse
    int nx = 0;                       // x-pass variants
    XKern xk[kMaxVariants];
};

// the compiled kernels of extent N, else the runtime-length path (nullptr
// outside [2, GEN_NMAX])
const FftEntry* find_entry(int N);
// the compiled kernels of extent N only
const FftEntry* find_compiled(int N);
// the runtime-length path for extent N (ose_generic.hip; nullptr outside
// [2, GEN_NMAX])
const FftEntry* generic_entry(int N);
constexpr int GEN_NMAX = 4096;
// smallest extent >= n compiled in several launch shapes (0 if none)
int next_supported(int n);
// 1: compiled in several launch shapes, 3: compiled in one, 2: the
// runtime-length path, 0: unsupported
int extent_kind(int n);

}  // namespace ose
