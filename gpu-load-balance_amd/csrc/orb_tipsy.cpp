// orb_tipsy.cpp -- Tipsy snapshot reader (host code).
//
// Layout (the published Tipsy format): a header of double time, int
// nbodies, ndim, nsph, ndark, nstar (and, in the standard XDR files, one int
// of padding to 32 bytes), then nsph gas records of 12 floats (mass, pos[3],
// vel[3], rho, temp, hsmooth, metals, phi), ndark dark records of 9 floats
// (mass, pos[3], vel[3], eps, phi) and nstar star records of 11 floats
// (mass, pos[3], vel[3], metals, tform, eps, phi).  Standard files are
// big-endian; native files are little-endian with the same 32-byte header.
#include <cstdio>
#include <cstring>
#include <memory>
#include <vector>

#include "orb.h"

namespace {

uint32_t bswap32(uint32_t v)
{
    return (v >> 24) | ((v >> 8) & 0xff00u) | ((v << 8) & 0xff0000u) | (v << 24);
}

uint64_t bswap64(uint64_t v)
{
    return ((uint64_t)bswap32((uint32_t)v) << 32) | bswap32((uint32_t)(v >> 32));
}

struct File {
    FILE *f = nullptr;
    ~File()
    {
        if (f) std::fclose(f);
    }
};

bool plausible(const orb_tipsy_header &h)
{
    return h.ndim == 3 && h.nsph >= 0 && h.ndark >= 0 && h.nstar >= 0 && h.nbodies >= 0 &&
           (int64_t)h.nsph + h.ndark + h.nstar == h.nbodies;
}

int read_header(FILE *f, orb_tipsy_header *h, long *data_offset)
{
    unsigned char raw[32];
    if (std::fread(raw, 1, 32, f) != 32) return ORB_ERR_ARG;
    for (int be = 1; be >= 0; --be) {
        orb_tipsy_header t;
        std::memset(&t, 0, sizeof t);
        uint64_t tb;
        std::memcpy(&tb, raw, 8);
        uint32_t w[5];
        std::memcpy(w, raw + 8, 20);
        if (be) {
            tb = bswap64(tb);
            for (auto &v : w) v = bswap32(v);
        }
        std::memcpy(&t.time, &tb, 8);
        t.nbodies = (int32_t)w[0];
        t.ndim = (int32_t)w[1];
        t.nsph = (int32_t)w[2];
        t.ndark = (int32_t)w[3];
        t.nstar = (int32_t)w[4];
        t.big_endian = be;
        if (plausible(t)) {
            *h = t;
            *data_offset = 32;
            return ORB_OK;
        }
    }
    return ORB_ERR_ARG;
}

}  // namespace

extern "C" {

int orb_read_tipsy(const char *path, orb_tipsy_header *hdr, float *x, float *y, float *z,
                   int64_t cap)
{
    if (!path || !hdr) return ORB_ERR_ARG;
    File fh;
    fh.f = std::fopen(path, "rb");
    if (!fh.f) return ORB_ERR_ARG;
    long off = 0;
    orb_tipsy_header h;
    if (read_header(fh.f, &h, &off) < 0) return ORB_ERR_ARG;
    *hdr = h;
    if (!x && !y && !z) return ORB_OK;
    if (!x || !y || !z || cap < h.nbodies) return ORB_ERR_ARG;
    // expected size: header + records
    const int64_t need = 32 + 48LL * h.nsph + 36LL * h.ndark + 44LL * h.nstar;
    if (std::fseek(fh.f, 0, SEEK_END) != 0) return ORB_ERR_ARG;
    if ((int64_t)std::ftell(fh.f) < need) return ORB_ERR_ARG;  // truncated
    if (std::fseek(fh.f, off, SEEK_SET) != 0) return ORB_ERR_ARG;
    const int nf[3] = {12, 9, 11};
    const int32_t cnt[3] = {h.nsph, h.ndark, h.nstar};
    std::vector<uint32_t> buf;
    int64_t k = 0;
    for (int s = 0; s < 3; ++s) {
        // stream in blocks of records
        const int64_t per = 1 << 16;
        for (int64_t done = 0; done < cnt[s]; done += per) {
            const int64_t m = cnt[s] - done < per ? cnt[s] - done : per;
            buf.resize((size_t)(m * nf[s]));
            if (std::fread(buf.data(), 4, buf.size(), fh.f) != buf.size()) return ORB_ERR_ARG;
            for (int64_t r = 0; r < m; ++r, ++k) {
                uint32_t p[3] = {buf[(size_t)(r * nf[s] + 1)], buf[(size_t)(r * nf[s] + 2)],
                                 buf[(size_t)(r * nf[s] + 3)]};
                if (h.big_endian)
                    for (auto &v : p) v = bswap32(v);
                std::memcpy(&x[k], &p[0], 4);
                std::memcpy(&y[k], &p[1], 4);
                std::memcpy(&z[k], &p[2], 4);
            }
        }
    }
    return ORB_OK;
}

int orb_init_tipsy(orb_ctx **ctx, const char *path, int32_t d, orb_meta_params params)
{
    if (!ctx) return ORB_ERR_ARG;
    *ctx = nullptr;
    orb_tipsy_header h;
    int rc = orb_read_tipsy(path, &h, nullptr, nullptr, nullptr, 0);
    if (rc < 0) return rc;
    std::unique_ptr<float[]> xs(new (std::nothrow) float[(size_t)h.nbodies * 3 + 1]);
    if (!xs) return ORB_ERR_NOMEM;
    float *x = xs.get(), *y = x + h.nbodies, *z = y + h.nbodies;
    if ((rc = orb_read_tipsy(path, &h, x, y, z, h.nbodies)) < 0) return rc;
    orb_ctx *c = nullptr;
    if ((rc = orb_init(&c, h.nbodies, d, false, params)) < 0) return rc;
    if ((rc = orb_set_particles(c, x, y, z, h.nbodies)) < 0) {
        orb_finalize(c);
        return rc;
    }
    *ctx = c;
    return ORB_OK;
}

}  // extern "C"
