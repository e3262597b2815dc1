// ose_host_io.cpp -- seismogram (raw, SU, SEG-Y) and snapshot files.
#include <algorithm>
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

// big-endian fields of SEG-Y files
void put_be32(unsigned char* h, int off, uint32_t v)
{
    h[off] = (unsigned char)(v >> 24);
    h[off + 1] = (unsigned char)(v >> 16);
    h[off + 2] = (unsigned char)(v >> 8);
    h[off + 3] = (unsigned char)v;
}
void put_be16(unsigned char* h, int off, uint16_t v)
{
    h[off] = (unsigned char)(v >> 8);
    h[off + 1] = (unsigned char)v;
}
uint32_t get_be32(const unsigned char* h, int off)
{
    return ((uint32_t)h[off] << 24) | ((uint32_t)h[off + 1] << 16) | ((uint32_t)h[off + 2] << 8) | h[off + 3];
}
uint16_t get_be16(const unsigned char* h, int off) { return (uint16_t)((h[off] << 8) | h[off + 1]); }

// The 240-byte trace header shared by SU (native byte order) and SEG-Y
// (big-endian); byte offsets of the SEG-Y rev 1 standard.
template <bool BE>
void trace_header(unsigned char* h, int r, int nt, double us, double gx, double gy, double gz, double src_x,
                  double src_y, double src_z, int shot_index)
{
    std::memset(h, 0, 240);
    auto i32 = [h](int off, int32_t v) {
        if (BE) put_be32(h, off, (uint32_t)v);
        else put<int32_t>(h, off, v);
    };
    auto i16 = [h](int off, int16_t v) {
        if (BE) put_be16(h, off, (uint16_t)v);
        else put<int16_t>(h, off, v);
    };
    i32(0, r + 1);                                 // tracl
    i32(4, r + 1);                                 // tracr
    i32(8, shot_index + 1);                        // fldr
    i32(12, r + 1);                                // tracf
    i32(16, shot_index + 1);                       // ep
    i32(20, r + 1);                                // cdp
    i16(28, 1);                                    // trid: seismic data
    i32(36, (int32_t)std::lround(std::hypot(gx - src_x, gy - src_y)));   // offset
    i32(40, (int32_t)std::lround(-gz * 100.0));    // gelev: receiver elevation (depth down)
    i32(48, (int32_t)std::lround(src_z * 100.0));  // sdepth: source depth
    i16(68, (int16_t)-100);                        // scalel: elevations / 100
    i16(70, (int16_t)-100);                        // scalco: coordinates / 100
    i32(72, (int32_t)std::lround(src_x * 100.0));
    i32(76, (int32_t)std::lround(src_y * 100.0));
    i32(80, (int32_t)std::lround(gx * 100.0));
    i32(84, (int32_t)std::lround(gy * 100.0));
    i16(88, 1);                                    // counit: length (m)
    i16(114, (int16_t)(uint16_t)nt);               // ns
    i16(116, (int16_t)(uint16_t)std::lround(us));  // dt (us)
}

// ASCII -> EBCDIC (code page 037) for the characters of the textual header
unsigned char ebcdic(char c)
{
    if (c >= 'a' && c <= 'z') c = (char)(c - 'a' + 'A');
    if (c >= '0' && c <= '9') return (unsigned char)(0xF0 + (c - '0'));
    if (c >= 'A' && c <= 'I') return (unsigned char)(0xC1 + (c - 'A'));
    if (c >= 'J' && c <= 'R') return (unsigned char)(0xD1 + (c - 'J'));
    if (c >= 'S' && c <= 'Z') return (unsigned char)(0xE2 + (c - 'S'));
    switch (c) {
    case '.': return 0x4B;
    case '(': return 0x4D;
    case '+': return 0x4E;
    case '-': return 0x60;
    case '/': return 0x61;
    case ',': return 0x6B;
    case '_': return 0x6D;
    case ')': return 0x5D;
    case ':': return 0x7A;
    case '=': return 0x7E;
    default: return 0x40;   // space
    }
}

// IBM System/360 single precision -> IEEE
float ibm_to_ieee(uint32_t b)
{
    const double sign = (b >> 31) ? -1.0 : 1.0;
    const int expo = (int)((b >> 24) & 0x7F) - 64;
    const double frac = (double)(b & 0x00FFFFFF) / 16777216.0;
    return (float)(sign * frac * std::pow(16.0, expo));
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
                  const double* rec_x, const double* rec_y, const double* rec_z, double src_x, double src_y,
                  double src_z,
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
        const double gx = rec_x ? rec_x[r] : 0.0, gy = rec_y ? rec_y[r] : 0.0, gz = rec_z ? rec_z[r] : 0.0;
        trace_header<false>(h, r, nt, us, gx, gy, gz, src_x, src_y, src_z, shot_index);
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
    size_t got;
    while ((got = std::fread(h, 1, sizeof h, f)) == sizeof h) {
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
    // a file that ends inside a trace header is cut short, not complete
    if (rc == OSE_OK && got != sizeof h && got != 0) rc = OSE_EINVAL;
    std::fclose(f);
    if (nrec) *nrec = n;
    if (nt) *nt = ns < 0 ? 0 : ns;
    if (dt) *dt = d;
    return rc;
}

int oseh_write_segy(const char* path, const float* traces, int nrec, int nt, double dt, const double* rec_x,
                    const double* rec_y, const double* rec_z, double src_x, double src_y, double src_z,
                    int shot_index)
{
    if (!path || !traces || nrec < 0 || nt < 0 || nt > 32767) return OSE_EINVAL;
    const double us = dt * 1e6;
    if (!(us > 0) || us > 32767.0) return OSE_EINVAL;
    FILE* f = std::fopen(path, "wb");
    if (!f) return OSE_EINVAL;
    // textual header: 40 card images of 80 EBCDIC characters
    unsigned char text[3200];
    char line[96];
    for (int c = 0; c < 40; ++c) {
        switch (c) {
        case 0: std::snprintf(line, sizeof line, "C 1 OSE ONE-STEP EXTRAPOLATION SEISMOGRAM, SHOT %d", shot_index + 1); break;
        case 1: std::snprintf(line, sizeof line, "C 2 TRACES %d SAMPLES %d INTERVAL %.0f US", nrec, nt, us); break;
        case 2: std::snprintf(line, sizeof line, "C 3 COORDINATES IN METRES TIMES 100 (SCALCO -100)"); break;
        case 38: std::snprintf(line, sizeof line, "C39 SEG Y REV1"); break;
        case 39: std::snprintf(line, sizeof line, "C40 END TEXTUAL HEADER"); break;
        default: std::snprintf(line, sizeof line, "C%2d", c + 1); break;
        }
        const size_t n = std::strlen(line);
        for (int k = 0; k < 80; ++k) text[c * 80 + k] = ebcdic(k < (int)n ? line[k] : ' ');
    }
    // binary header (file bytes 3201-3600)
    unsigned char bin[400];
    std::memset(bin, 0, sizeof bin);
    put_be32(bin, 0, 1);                                     // job id
    put_be32(bin, 4, 1);                                     // line number
    put_be32(bin, 8, 1);                                     // reel number
    put_be16(bin, 12, (uint16_t)std::min(nrec, 32767));      // traces per ensemble
    put_be16(bin, 16, (uint16_t)std::lround(us));            // sample interval
    put_be16(bin, 18, (uint16_t)std::lround(us));            // original sample interval
    put_be16(bin, 20, (uint16_t)nt);                         // samples per trace
    put_be16(bin, 22, (uint16_t)nt);
    put_be16(bin, 24, 5);                                    // format: 4-byte IEEE float
    put_be16(bin, 26, 1);                                    // ensemble fold
    put_be16(bin, 28, 1);                                    // sorting: as recorded
    put_be16(bin, 54, 1);                                    // measurement system: metres
    put_be16(bin, 300, 0x0100);                              // SEG Y revision 1.0
    put_be16(bin, 302, 1);                                   // fixed trace length
    bool ok = std::fwrite(text, 1, sizeof text, f) == sizeof text && std::fwrite(bin, 1, sizeof bin, f) == sizeof bin;
    unsigned char h[240];
    std::vector<unsigned char> data((size_t)nt * 4);
    for (int r = 0; r < nrec && ok; ++r) {
        const double gx = rec_x ? rec_x[r] : 0.0, gy = rec_y ? rec_y[r] : 0.0, gz = rec_z ? rec_z[r] : 0.0;
        trace_header<true>(h, r, nt, us, gx, gy, gz, src_x, src_y, src_z, shot_index);
        for (int i = 0; i < nt; ++i) {
            uint32_t u;
            std::memcpy(&u, traces + (size_t)r * nt + i, 4);
            put_be32(data.data(), 4 * i, u);
        }
        ok = std::fwrite(h, 1, sizeof h, f) == sizeof h && std::fwrite(data.data(), 1, data.size(), f) == data.size();
    }
    const int rc = std::fclose(f);
    return (ok && rc == 0) ? OSE_OK : OSE_EINVAL;
}

int oseh_read_segy(const char* path, float* traces, size_t cap, int* nrec, int* nt, double* dt)
{
    if (!path) return OSE_EINVAL;
    FILE* f = std::fopen(path, "rb");
    if (!f) return OSE_EINVAL;
    unsigned char head[3600];
    if (std::fread(head, 1, sizeof head, f) != sizeof head) {
        std::fclose(f);
        return OSE_EINVAL;
    }
    const int fmt = get_be16(head, 3224);
    const int ns = get_be16(head, 3220);
    const double d = get_be16(head, 3216) * 1e-6;
    if ((fmt != 1 && fmt != 5) || ns <= 0) {
        std::fclose(f);
        return OSE_EINVAL;
    }
    unsigned char h[240];
    std::vector<unsigned char> buf((size_t)ns * 4);
    int n = 0, rc = OSE_OK;
    size_t used = 0, got;
    while ((got = std::fread(h, 1, sizeof h, f)) == sizeof h) {
        if (get_be16(h, 114) != ns || std::fread(buf.data(), 1, buf.size(), f) != buf.size()) {
            rc = OSE_EINVAL;
            break;
        }
        if (traces) {
            if (used + (size_t)ns > cap) {
                rc = OSE_EINVAL;
                break;
            }
            for (int i = 0; i < ns; ++i) {
                const uint32_t u = get_be32(buf.data(), 4 * i);
                float v;
                if (fmt == 5) std::memcpy(&v, &u, 4);
                else v = ibm_to_ieee(u);
                traces[used + i] = v;
            }
        }
        used += (size_t)ns;
        ++n;
    }
    // a file that ends inside a trace header is cut short, not complete
    if (rc == OSE_OK && got != sizeof h && got != 0) rc = OSE_EINVAL;
    std::fclose(f);
    if (nrec) *nrec = n;
    if (nt) *nt = ns;
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

// ---- encoded snapshots (the layouts of ose_set_snapshots) -----------------

static int snap_vbytes(int format) { return format == OSE_SNAP_F32 ? 4 : format == OSE_SNAP_Q8 ? 1 : 2; }
static bool snap_scaled(int format) { return format == OSE_SNAP_Q16 || format == OSE_SNAP_Q8; }

size_t oseh_snapshot_lane_bytes(int format, int nx, int ny, int nz)
{
    if (format < OSE_SNAP_F32 || format > OSE_SNAP_Q8 || nx <= 0 || ny <= 0 || nz <= 0) return 0;
    const size_t cells = (size_t)nx * ny * nz;
    return (snap_scaled(format) ? sizeof(float) * (size_t)ny * nz : 0) + cells * snap_vbytes(format);
}

int oseh_snapshot_lane(int format, int nl, int lane, int nx, int ny, int nz, const void* data, void* out)
{
    if (!data || !out || nl < 1 || nl > 2 || lane < 0 || lane >= nl) return OSE_EINVAL;
    const size_t lb = oseh_snapshot_lane_bytes(format, nx, ny, nz);
    if (!lb) return OSE_EINVAL;
    const size_t cells = (size_t)nx * ny * nz, lines = (size_t)ny * nz;
    const char* d = (const char*)data;
    char* o = (char*)out;
    size_t vo = 0;
    if (snap_scaled(format)) {
        std::memcpy(o, d + sizeof(float) * lines * lane, sizeof(float) * lines);
        o += sizeof(float) * lines;
        vo = sizeof(float) * lines * nl;
    }
    const size_t vb = cells * snap_vbytes(format);
    std::memcpy(o, d + vo + vb * lane, vb);
    return OSE_OK;
}

int oseh_decode_snapshot(int format, int nx, int ny, int nz, const void* lane_data, float* out)
{
    if (!lane_data || !out || !oseh_snapshot_lane_bytes(format, nx, ny, nz)) return OSE_EINVAL;
    const size_t cells = (size_t)nx * ny * nz, lines = (size_t)ny * nz;
    const char* d = (const char*)lane_data;
    if (format == OSE_SNAP_F32) {
        std::memcpy(out, d, sizeof(float) * cells);
    } else if (format == OSE_SNAP_BF16) {
        for (size_t i = 0; i < cells; ++i) {
            uint16_t h;
            std::memcpy(&h, d + 2 * i, 2);
            const uint32_t b = (uint32_t)h << 16;
            std::memcpy(out + i, &b, 4);
        }
    } else {
        const float* sc = (const float*)d;
        const float qmax = format == OSE_SNAP_Q8 ? 127.f : 32767.f;
        const char* v = d + sizeof(float) * lines;
        for (size_t l = 0; l < lines; ++l) {
            const float step = sc[l] / qmax;
            for (int x = 0; x < nx; ++x) {
                const size_t i = l * nx + x;
                float q;
                if (format == OSE_SNAP_Q8) {
                    q = (float)(int8_t)v[i];
                } else {
                    int16_t t;
                    std::memcpy(&t, v + 2 * i, 2);
                    q = (float)t;
                }
                out[i] = q * step;
            }
        }
    }
    return OSE_OK;
}

// .osnap: 64-byte little-endian header, then one lane's encoded data
//   0 "OSESNAP1"  8 format  12 lane (0 P, 1 Q)  16 step  20 shot
//   24 x0 y0 z0  36 nx ny nz  48 u64 payload bytes  56 reserved
int oseh_write_osnap(const char* path, int format, int lane, int step, int shot, const int* box,
                     const void* lane_data)
{
    if (!path || !box || !lane_data) return OSE_EINVAL;
    const size_t pb = oseh_snapshot_lane_bytes(format, box[3], box[4], box[5]);
    if (!pb) return OSE_EINVAL;
    unsigned char h[64] = {};
    std::memcpy(h, "OSESNAP1", 8);
    const int32_t f[10] = {format, lane, step, shot, box[0], box[1], box[2], box[3], box[4], box[5]};
    std::memcpy(h + 8, f, sizeof f);
    const uint64_t n = pb;
    std::memcpy(h + 48, &n, 8);
    FILE* fp = std::fopen(path, "wb");
    if (!fp) return OSE_EINVAL;
    const bool ok = std::fwrite(h, 1, 64, fp) == 64 && std::fwrite(lane_data, 1, pb, fp) == pb;
    const int rc = std::fclose(fp);
    return (ok && rc == 0) ? OSE_OK : OSE_EINVAL;
}

int oseh_read_osnap(const char* path, float* out, size_t cap, int* meta)
{
    if (!path || !meta) return OSE_EINVAL;
    FILE* fp = std::fopen(path, "rb");
    if (!fp) return OSE_EINVAL;
    unsigned char h[64];
    int rc = OSE_OK;
    int32_t f[10] = {};
    uint64_t n = 0;
    if (std::fread(h, 1, 64, fp) != 64 || std::memcmp(h, "OSESNAP1", 8) != 0) rc = OSE_EINVAL;
    if (!rc) {
        std::memcpy(f, h + 8, sizeof f);
        std::memcpy(&n, h + 48, 8);
        if (n != oseh_snapshot_lane_bytes(f[0], f[7], f[8], f[9]) || n == 0) rc = OSE_EINVAL;
    }
    if (!rc) {
        std::memcpy(meta, f, sizeof f);
        const size_t cells = (size_t)f[7] * f[8] * f[9];
        if (out) {
            std::vector<char> buf(n);
            if (cap < cells || std::fread(buf.data(), 1, n, fp) != n || std::fgetc(fp) != EOF) rc = OSE_EINVAL;
            else rc = oseh_decode_snapshot(f[0], f[7], f[8], f[9], buf.data(), out);
        }
    }
    std::fclose(fp);
    return rc;
}

}  // extern "C"
