// Snapshot files in the reference's ASCII column format, written and parsed
// on host threads.
//
// Layout (FDPS writeParticleAscii with FileHeader::writeAscii,
// src/class.h:45-58, and RealPtcl::writeAscii, src/class.h:129-135):
//   "%e\n"  time
//   "%d\n"  N
//   N lines: id \t tag \t mass \t x \t y \t z \t vx \t vy \t vz \t dens \t snds \t pres
//            ("%ld" for id and tag, "%.16lf" for the rest)
// 2D writes z = vz = 0 and the sound speed; 3D writes snds = 0.
#include <cerrno>
#include <cinttypes>
#include <cstdlib>
#include <cstring>
#include <memory>

#include "common.hpp"

namespace wcsph {
namespace {

struct FileCloser {
    void operator()(FILE* f) const {
        if (f) std::fclose(f);
    }
};
using File = std::unique_ptr<FILE, FileCloser>;

File open_file(const char* path, const char* mode) {
    WCSPH_REQUIRE(path != nullptr, WCSPH_EINVAL, "null path");
    FILE* f = std::fopen(path, mode);
    WCSPH_REQUIRE(f != nullptr, WCSPH_EIO, std::string("cannot open ") + path + ": " + std::strerror(errno));
    return File(f);
}

constexpr const char* kLineFmt =
    "%" PRId64 "\t%" PRId64 "\t%.16f\t%.16f\t%.16f\t%.16f\t%.16f\t%.16f\t%.16f\t%.16f\t%.16f\t%.16f\n";

// Appends one particle line (src/class.h:131-133 in 2D, :133 in 3D).
void append_line(std::string& out, int64_t id, int64_t tag, const double* v) {
    char buf[512];
    const int len = std::snprintf(buf, sizeof buf, kLineFmt, id, tag, v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7],
                                  v[8], v[9]);
    if (len < (int)sizeof buf) {
        out.append(buf, (size_t)len);
        return;
    }
    // values beyond ~1e300 print hundreds of digits
    std::vector<char> big((size_t)len + 1);
    std::snprintf(big.data(), big.size(), kLineFmt, id, tag, v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7], v[8],
                  v[9]);
    out.append(big.data(), (size_t)len);
}

}  // namespace
}  // namespace wcsph

using namespace wcsph;

extern "C" {

int wcsph_snapshot_write_ascii(const char* path, int32_t dim, int64_t n, double time, const int64_t* id,
                               const int64_t* tag, const double* mass, const double* pos, const double* vel,
                               const double* dens, const double* snds, const double* pres) {
    return guarded([&] {
        WCSPH_REQUIRE(dim == 2 || dim == 3, WCSPH_EINVAL, "dim must be 2 or 3");
        WCSPH_REQUIRE(n >= 0 && n <= INT32_MAX, WCSPH_EINVAL, "n must be in [0, 2^31) (the header's %d)");
        WCSPH_REQUIRE(n == 0 || (mass && pos && vel && dens), WCSPH_EINVAL, "null mass, pos, vel or dens");
        File f = open_file(path, "wb");
        WCSPH_REQUIRE(std::fprintf(f.get(), "%e\n%d\n", time, (int)n) > 0, WCSPH_EIO, "write failed");
        // lines are formatted in parallel in chunks and written in order
        const int64_t chunk = 8192;
        const int64_t nchunk = ceil_div(n, chunk);
        const int64_t batch = 64;
        std::vector<std::string> text((size_t)std::min(nchunk, batch));
        for (int64_t c0 = 0; c0 < nchunk; c0 += batch) {
            const int64_t nc = std::min(batch, nchunk - c0);
            parallel_for(
                nc,
                [&](int64_t k) {
                    std::string& out = text[(size_t)k];
                    out.clear();
                    const int64_t b = (c0 + k) * chunk, e = std::min(n, b + chunk);
                    out.reserve((size_t)(e - b) * 240);
                    for (int64_t i = b; i < e; ++i) {
                        double v[10];
                        v[0] = mass[i];
                        v[1] = pos[i];
                        v[2] = pos[n + i];
                        v[3] = dim == 3 ? pos[2 * n + i] : 0.0;
                        v[4] = vel[i];
                        v[5] = vel[n + i];
                        v[6] = dim == 3 ? vel[2 * n + i] : 0.0;
                        v[7] = dens[i];
                        v[8] = (dim == 2 && snds) ? snds[i] : 0.0;
                        v[9] = pres ? pres[i] : 0.0;
                        append_line(out, id ? id[i] : i, tag ? tag[i] : 0, v);
                    }
                },
                nc < 2 ? 1 : 0);
            for (int64_t k = 0; k < nc; ++k) {
                const std::string& out = text[(size_t)k];
                WCSPH_REQUIRE(std::fwrite(out.data(), 1, out.size(), f.get()) == out.size(), WCSPH_EIO,
                              std::string("write failed: ") + path);
            }
        }
        WCSPH_REQUIRE(std::fflush(f.get()) == 0, WCSPH_EIO, std::string("write failed: ") + path);
    });
}

int wcsph_snapshot_read_ascii(const char* path, int32_t dim, int64_t* n, double* time, int64_t* id, int64_t* tag,
                              double* mass, double* pos, double* vel, double* dens, double* snds, double* pres) {
    return guarded([&] {
        WCSPH_REQUIRE(dim == 2 || dim == 3, WCSPH_EINVAL, "dim must be 2 or 3");
        WCSPH_REQUIRE(n != nullptr, WCSPH_EINVAL, "null n");
        File f = open_file(path, "rb");
        WCSPH_REQUIRE(std::fseek(f.get(), 0, SEEK_END) == 0, WCSPH_EIO, "cannot seek");
        const long size = std::ftell(f.get());
        WCSPH_REQUIRE(size >= 0, WCSPH_EIO, "cannot size the file");
        std::rewind(f.get());
        std::vector<char> buf((size_t)size + 1);
        WCSPH_REQUIRE(std::fread(buf.data(), 1, (size_t)size, f.get()) == (size_t)size, WCSPH_EIO, "read failed");
        buf[(size_t)size] = '\0';
        char* p = buf.data();
        char* end = nullptr;
        const double t = std::strtod(p, &end);
        WCSPH_REQUIRE(end != p, WCSPH_EIO, "snapshot header: no time");
        p = end;
        const long long cnt = std::strtoll(p, &end, 10);
        WCSPH_REQUIRE(end != p && cnt >= 0, WCSPH_EIO, "snapshot header: no particle count");
        p = end;
        const bool only_header = !(id || tag || mass || pos || vel || dens || snds || pres);
        if (!only_header) WCSPH_REQUIRE(*n == cnt, WCSPH_EINVAL, "*n differs from the file's particle count");
        *n = cnt;
        if (time) *time = t;
        if (only_header) return;
        // line starts (after the header's line break)
        while (*p == '\r' || *p == '\n') ++p;
        std::vector<char*> line((size_t)cnt);
        for (long long i = 0; i < cnt; ++i) {
            WCSPH_REQUIRE(*p != '\0', WCSPH_EIO,
                          "snapshot truncated: " + std::to_string(i) + " of " + std::to_string(cnt) + " lines");
            line[(size_t)i] = p;
            char* nl = std::strchr(p, '\n');
            if (!nl) {
                p += std::strlen(p);
            } else {
                *nl = '\0';
                p = nl + 1;
            }
        }
        std::vector<uint8_t> bad((size_t)cnt, 0);
        parallel_for(cnt, [&](int64_t i) {
            char* q = line[(size_t)i];
            char* e = nullptr;
            const long long a = std::strtoll(q, &e, 10);
            bool ok = e != q;
            q = e;
            const long long b = std::strtoll(q, &e, 10);
            ok = ok && e != q;
            q = e;
            double v[10];
            for (int k = 0; k < 10; ++k) {
                v[k] = std::strtod(q, &e);
                ok = ok && e != q;
                q = e;
            }
            if (!ok) {
                bad[(size_t)i] = 1;
                return;
            }
            const int64_t nn = cnt;
            if (id) id[i] = a;
            if (tag) tag[i] = b;
            if (mass) mass[i] = v[0];
            if (pos) {
                pos[i] = v[1];
                pos[nn + i] = v[2];
                if (dim == 3) pos[2 * nn + i] = v[3];
            }
            if (vel) {
                vel[i] = v[4];
                vel[nn + i] = v[5];
                if (dim == 3) vel[2 * nn + i] = v[6];
            }
            if (dens) dens[i] = v[7];
            if (snds) snds[i] = v[8];
            if (pres) pres[i] = v[9];
        });
        for (long long i = 0; i < cnt; ++i)
            WCSPH_REQUIRE(!bad[(size_t)i], WCSPH_EIO,
                          "snapshot line " + std::to_string(i + 3) + ": expected 12 numeric columns");
    });
}

}  // extern "C"
