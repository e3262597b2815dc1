// ose_host_io.cpp -- seismogram (raw and SU) and snapshot files.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "ose_host.h"

namespace {

template <typename T>
void put(unsigned char* h, int off, T v)
{
    std::memcpy(h + off, &v, sizeof(T));
}
template <typename T>
T get(const unsigned char* h, int off)
{
    T v;
    std::memcpy(&v, h + off, sizeof(T));
    return v;
}

}  // namespace

extern "C" {

int oseh_write_raw(const char* path, const float* data, size_t n)
{
    if (!path || (!data && n)) return OSE_EINVAL;
    FILE* f = std::fopen(path, "wb");
    if (!f) return OSE_EINVAL;
    const size_t w = std::fwrite(data, sizeof(float), n, f);
    const int rc = std::fclose(f);
    return (w == n && rc == 0) ? OSE_OK : OSE_EINVAL;
}

int oseh_write_su(const char* path, const float* traces, int nrec, int nt, double dt,
                  const double* rec_x, const double* rec_y, double src_x, double src_y,
                  int shot_index)
{
    if (!path || !traces || nrec < 0 || nt < 0 || nt > 65535) return OSE_EINVAL;
    const double us = dt * 1e6;
    if (!(us > 0) || us > 65535.0) return OSE_EINVAL;
    FILE* f = std::fopen(path, "wb");
    if (!f) return OSE_EINVAL;
    unsigned char h[240];
    bool ok = true;
    for (int r = 0; r < nrec && ok; ++r) {
        std::memset(h, 0, sizeof h);
        put<int32_t>(h, 0, r + 1);                 // tracl
        put<int32_t>(h, 4, r + 1);                 // tracr
        put<int32_t>(h, 8, shot_index + 1);        // fldr
        put<int32_t>(h, 12, r + 1);                // tracf
        put<int32_t>(h, 16, shot_index + 1);       // ep
        put<int32_t>(h, 20, r + 1);                // cdp
        put<int16_t>(h, 28, 1);                    // trid: seismic data
        const double gx = rec_x ? rec_x[r] : 0.0, gy = rec_y ? rec_y[r] : 0.0;
        put<int32_t>(h, 36, (int32_t)std::lround(std::hypot(gx - src_x, gy - src_y)));   // offset
        put<int16_t>(h, 70, (int16_t)-100);        // scalco: coordinates / 100
        put<int32_t>(h, 72, (int32_t)std::lround(src_x * 100.0));
        put<int32_t>(h, 76, (int32_t)std::lround(src_y * 100.0));
        put<int32_t>(h, 80, (int32_t)std::lround(gx * 100.0));
        put<int32_t>(h, 84, (int32_t)std::lround(gy * 100.0));
        put<int16_t>(h, 88, 1);                    // counit: length (m)
        put<uint16_t>(h, 114, (uint16_t)nt);       // ns
        put<uint16_t>(h, 116, (uint16_t)std::lround(us));   // dt (us)
        ok = std::fwrite(h, 1, sizeof h, f) == sizeof h &&
             std::fwrite(traces + (size_t)r * nt, sizeof(float), (size_t)nt, f) == (size_t)nt;
    }
    const int rc = std::fclose(f);
    return (ok && rc == 0) ? OSE_OK : OSE_EINVAL;
}

int oseh_read_su(const char* path, float* traces, size_t cap, int* nrec, int* nt, double* dt)
{
    if (!path) return OSE_EINVAL;
    FILE* f = std::fopen(path, "rb");
    if (!f) return OSE_EINVAL;
    unsigned char h[240];
    int n = 0, ns = -1;
    double d = 0.0;
    size_t used = 0;
    int rc = OSE_OK;
    while (std::fread(h, 1, sizeof h, f) == sizeof h) {
        const int tns = get<uint16_t>(h, 114);
        if (ns < 0) {
            ns = tns;
            d = get<uint16_t>(h, 116) * 1e-6;
        } else if (tns != ns) {
            rc = OSE_EINVAL;
            break;
        }
        if (traces && used + (size_t)ns > cap) {
            rc = OSE_EINVAL;
            break;
        }
        std::vector<float> buf((size_t)ns);
        if (std::fread(buf.data(), sizeof(float), (size_t)ns, f) != (size_t)ns) {
            rc = OSE_EINVAL;
            break;
        }
        if (traces) std::memcpy(traces + used, buf.data(), sizeof(float) * (size_t)ns);
        used += (size_t)ns;
        ++n;
    }
    std::fclose(f);
    if (nrec) *nrec = n;
    if (nt) *nt = ns < 0 ? 0 : ns;
    if (dt) *dt = d;
    return rc;
}

int oseh_write_snapshot(const char* path, const float* f_zyx, int nxb, int nyb, int nzb)
{
    // write_3dfile_from_1darray_ci (ose_3d_lib.cu:364-380): loops y, x, z
    if (!path || !f_zyx) return OSE_EINVAL;
    FILE* f = std::fopen(path, "wb");
    if (!f) return OSE_EINVAL;
    std::vector<float> col((size_t)nzb);
    bool ok = true;
    for (int y = 0; y < nyb && ok; ++y)
        for (int x = 0; x < nxb && ok; ++x) {
            for (int z = 0; z < nzb; ++z) col[z] = f_zyx[((size_t)z * nyb + y) * nxb + x];
            ok = std::fwrite(col.data(), sizeof(float), col.size(), f) == col.size();
        }
    const int rc = std::fclose(f);
    return (ok && rc == 0) ? OSE_OK : OSE_EINVAL;
}

}  // extern "C"
