// gt_ingest.hip -- movie files for batch_file / parse_ets (host code).
//
// TIFF (decoderiter.rs:9-71 reads pages with the `tiff` crate and keeps
// u8/u16 grey pages): classic and BigTIFF, either byte order, every page of
// one shape, strips (no tiles), one sample per pixel, 8 or 16 bit unsigned,
// compression none, LZW, Deflate or PackBits, optional horizontal predictor.
//
// ETS (decoderiter.rs:77-260, MinimalETSParser): the SIS header gives the
// number of dimensions, the ETS header offset and the chunk table; the ETS
// header gives the data type and the frame size (dims[0], dims[1]); every
// chunk record holds its coordinates (frame = coordinate 2, channel =
// coordinate 4) and the file offset of its pixels.  Frames of a channel are
// placed in the order the chunk table gives them, with the reference's rule
// for repeated frame coordinates (a chunk whose frame coordinate is the last
// one filled is appended after it); missing frames are skipped.  Pixels are
// little-endian uint16 (dims[0] * dims[1] values per frame); other data
// types are rejected, as the reference's reader only decodes uint16.
//
// Reads use pread, so one gt_file serves several threads.
#include <fcntl.h>
#include <unistd.h>
#include <zlib.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <new>
#include <string>
#include <vector>

#include "gpu_tracking.h"
#include "gt_internal.h"

struct gt_file {
    int fd = -1;
    int kind = 0;            // 0 TIFF, 1 ETS
    bool big_endian = false;
    uint32_t h = 0, w = 0;   // frame shape (ETS: rows dims[1], columns dims[0])
    int dtype = 0;           // 0 u16, 2 u8
    uint64_t file_size = 0;
    // TIFF pages
    struct Page {
        std::vector<uint64_t> off, cnt;
        uint32_t rows_per_strip = 0;
        uint32_t compression = 1, predictor = 1;
    };
    std::vector<Page> pages;
    // ETS frames of the selected channel
    std::vector<uint64_t> ets_off;
    std::vector<uint32_t> channels;
    ~gt_file() {
        if (fd >= 0) close(fd);
    }
};

namespace gt {
gt_status set_last_error(gt_status s, const std::string& msg);
}

namespace {

using gt::set_last_error;

bool pread_all(int fd, void* dst, size_t n, uint64_t off) {
    char* p = (char*)dst;
    while (n) {
        const ssize_t r = pread(fd, p, n, (off_t)off);
        if (r <= 0) return false;
        p += r;
        n -= (size_t)r;
        off += (uint64_t)r;
    }
    return true;
}

struct Reader {
    const gt_file& f;
    std::string err;
    bool rd(void* dst, size_t n, uint64_t off) {
        if (off > f.file_size || n > f.file_size - off) {
            err = "read beyond the end of the file";
            return false;
        }
        if (!pread_all(f.fd, dst, n, off)) {
            err = "read failed";
            return false;
        }
        return true;
    }
    uint64_t swap(uint64_t v, int n) const {
        if (!f.big_endian) return v;
        uint64_t r = 0;
        for (int i = 0; i < n; ++i) r |= ((v >> (8 * i)) & 0xff) << (8 * (n - 1 - i));
        return r;
    }
    bool u(uint64_t off, int n, uint64_t& v) {
        uint64_t x = 0;
        if (!rd(&x, (size_t)n, off)) return false;
        v = swap(x, n);
        return true;
    }
};

int type_size(uint32_t t) {
    switch (t) {
        case 1: case 2: case 6: case 7: return 1;
        case 3: case 8: return 2;
        case 4: case 9: case 11: case 13: return 4;
        case 5: case 10: case 12: case 16: case 17: case 18: return 8;
        default: return 0;
    }
}

// ---- TIFF ------------------------------------------------------------------

gt_status open_tiff(gt_file& f) {
    Reader R{f, {}};
    char bo[2];
    if (!R.rd(bo, 2, 0)) return set_last_error(GT_E_INVALID, "TIFF: " + R.err);
    if (bo[0] == 'I' && bo[1] == 'I') f.big_endian = false;
    else if (bo[0] == 'M' && bo[1] == 'M') f.big_endian = true;
    else return set_last_error(GT_E_INVALID, "not a TIFF file");
    uint64_t magic, ifd;
    if (!R.u(2, 2, magic)) return set_last_error(GT_E_INVALID, "TIFF: " + R.err);
    const bool big = magic == 43;
    if (magic != 42 && !big) return set_last_error(GT_E_INVALID, "not a TIFF file");
    if (!R.u(big ? 8 : 4, big ? 8 : 4, ifd)) return set_last_error(GT_E_INVALID, "TIFF: " + R.err);
    std::vector<uint64_t> seen;   // IFD offsets visited (a chain must not loop)
    while (ifd) {
        if (std::find(seen.begin(), seen.end(), ifd) != seen.end() || seen.size() > (1u << 24))
            return set_last_error(GT_E_INVALID, "TIFF: IFD chain does not end");
        seen.push_back(ifd);
        uint64_t n;
        if (!R.u(ifd, big ? 8 : 2, n)) return set_last_error(GT_E_INVALID, "TIFF: " + R.err);
        const uint64_t esz = big ? 20 : 12;
        if (n > 4096) return set_last_error(GT_E_INVALID, "TIFF: implausible IFD entry count");
        std::map<uint32_t, std::vector<uint64_t>> tags;
        for (uint64_t i = 0; i < n; ++i) {
            const uint64_t e = ifd + (big ? 8 : 2) + i * esz;
            uint64_t tag, typ, cnt;
            if (!R.u(e, 2, tag) || !R.u(e + 2, 2, typ) || !R.u(e + 4, big ? 8 : 4, cnt))
                return set_last_error(GT_E_INVALID, "TIFF: " + R.err);
            const int ts = type_size((uint32_t)typ);
            if (!(typ == 1 || typ == 3 || typ == 4 || typ == 16 || typ == 13 || typ == 18))
                continue;   // only integer tags are needed
            if (cnt > (1ull << 26)) return set_last_error(GT_E_INVALID, "TIFF: implausible tag count");
            const uint64_t inl = big ? 8 : 4;
            uint64_t at = e + (big ? 12 : 8);
            if (cnt * (uint64_t)ts > inl) {
                if (!R.u(at, (int)inl, at)) return set_last_error(GT_E_INVALID, "TIFF: " + R.err);
            }
            std::vector<uint64_t> vals((size_t)cnt);
            std::vector<uint8_t> raw((size_t)(cnt * ts));
            if (cnt && !R.rd(raw.data(), raw.size(), at))
                return set_last_error(GT_E_INVALID, "TIFF: " + R.err);
            for (uint64_t k = 0; k < cnt; ++k) {
                uint64_t v = 0;
                std::memcpy(&v, raw.data() + k * ts, (size_t)ts);
                vals[(size_t)k] = R.swap(v, ts);
            }
            tags[(uint32_t)tag] = std::move(vals);
        }
        auto one = [&](uint32_t t, uint64_t d) {
            auto it = tags.find(t);
            return (it == tags.end() || it->second.empty()) ? d : it->second[0];
        };
        const uint64_t w = one(256, 0), h = one(257, 0);
        const uint64_t bps = one(258, 1), comp = one(259, 1), spp = one(277, 1);
        const uint64_t fmt = one(339, 1), pred = one(317, 1), planar = one(284, 1);
        if (!w || !h || w > (1u << 20) || h > (1u << 20))
            return set_last_error(GT_E_INVALID, "TIFF: bad page size");
        if (tags.count(322) || tags.count(324))
            return set_last_error(GT_E_NOT_IMPLEMENTED, "TIFF: tiled pages are not supported");
        if (spp != 1 || planar != 1 || fmt != 1 || (bps != 8 && bps != 16))
            return set_last_error(GT_E_NOT_IMPLEMENTED,
                                  "TIFF: only 8/16-bit unsigned single-sample pages (the reference "
                                  "reads u8 and u16 grey pages)");
        if (comp != 1 && comp != 5 && comp != 8 && comp != 32946 && comp != 32773)
            return set_last_error(GT_E_NOT_IMPLEMENTED,
                                  "TIFF: compression " + std::to_string(comp) + " is not supported");
        if (pred != 1 && pred != 2) return set_last_error(GT_E_NOT_IMPLEMENTED, "TIFF: predictor not supported");
        const int dt = bps == 16 ? 0 : 2;
        if (f.pages.empty()) {
            f.h = (uint32_t)h;
            f.w = (uint32_t)w;
            f.dtype = dt;
        } else if (f.h != h || f.w != w || f.dtype != dt) {
            return set_last_error(GT_E_INVALID, "TIFF: pages of different shapes or depths");
        }
        gt_file::Page pg;
        auto so = tags.find(273), sc = tags.find(279);
        if (so == tags.end() || sc == tags.end() || so->second.size() != sc->second.size())
            return set_last_error(GT_E_INVALID, "TIFF: missing or inconsistent strip tags");
        pg.off = so->second;
        pg.cnt = sc->second;
        pg.rows_per_strip = (uint32_t)std::min<uint64_t>(one(278, h), h);
        if (pg.rows_per_strip == 0) pg.rows_per_strip = (uint32_t)h;
        if (pg.off.size() != (h + pg.rows_per_strip - 1) / pg.rows_per_strip)
            return set_last_error(GT_E_INVALID, "TIFF: strip count does not match the page size");
        for (size_t k = 0; k < pg.off.size(); ++k)
            if (pg.off[k] > f.file_size || pg.cnt[k] > f.file_size - pg.off[k])
                return set_last_error(GT_E_INVALID, "TIFF: strip beyond the end of the file");
        pg.compression = (uint32_t)comp;
        pg.predictor = (uint32_t)pred;
        f.pages.push_back(std::move(pg));
        const uint64_t nx = ifd + (big ? 8 : 2) + n * esz;
        if (!R.u(nx, big ? 8 : 4, ifd)) return set_last_error(GT_E_INVALID, "TIFF: " + R.err);
    }
    if (f.pages.empty()) return set_last_error(GT_E_INVALID, "TIFF: no pages");
    return GT_OK;
}

// PackBits (TIFF 6.0 section 9).
bool unpackbits(const uint8_t* s, size_t n, uint8_t* d, size_t m) {
    size_t i = 0, o = 0;
    while (i < n && o < m) {
        const int8_t c = (int8_t)s[i++];
        if (c >= 0) {
            const size_t k = (size_t)c + 1;
            if (i + k > n || o + k > m) return false;
            std::memcpy(d + o, s + i, k);
            i += k;
            o += k;
        } else if (c != -128) {
            const size_t k = (size_t)(1 - c);
            if (i >= n || o + k > m) return false;
            std::memset(d + o, s[i++], k);
            o += k;
        }
    }
    return o == m;
}

// TIFF LZW (TIFF 6.0 section 13): MSB-first codes of 9..12 bits, Clear 256,
// EndOfInformation 257, the code width grows one code early.
bool unlzw(const uint8_t* s, size_t n, uint8_t* d, size_t m) {
    std::vector<uint32_t> prefix(4096);
    std::vector<uint8_t> suffix(4096), first(4096);
    std::vector<uint16_t> len(4096);
    for (int i = 0; i < 256; ++i) {
        prefix[i] = 0xffffffffu;
        suffix[i] = (uint8_t)i;
        first[i] = (uint8_t)i;
        len[i] = 1;
    }
    uint64_t bitpos = 0;
    const uint64_t nbits = (uint64_t)n * 8;
    int width = 9;
    uint32_t next = 258, old = 0xffffffffu;
    size_t o = 0;
    auto emit = [&](uint32_t code) -> bool {
        const size_t L = len[code];
        if (o + L > m) return false;
        uint32_t c = code;
        for (size_t k = L; k-- > 0;) {
            d[o + k] = suffix[c];
            c = prefix[c];
        }
        o += L;
        return true;
    };
    while (bitpos + width <= nbits) {
        uint32_t code = 0;
        for (int b = 0; b < width; ++b, ++bitpos)
            code = (code << 1) | ((s[bitpos >> 3] >> (7 - (bitpos & 7))) & 1u);
        if (code == 257) break;
        if (code == 256) {
            next = 258;
            width = 9;
            old = 0xffffffffu;
            continue;
        }
        if (old == 0xffffffffu) {
            if (code > 255 || !emit(code)) return false;
            old = code;
            continue;
        }
        if (code < next) {
            if (!emit(code)) return false;
            if (next < 4096) {
                prefix[next] = old;
                suffix[next] = first[code];
                first[next] = first[old];
                len[next] = (uint16_t)(len[old] + 1);
                ++next;
            }
        } else if (code == next && next < 4096) {
            prefix[next] = old;
            suffix[next] = first[old];
            first[next] = first[old];
            len[next] = (uint16_t)(len[old] + 1);
            ++next;
            if (!emit(code)) return false;
        } else {
            return false;
        }
        old = code;
        if (next + 1 >= (1u << width) && width < 12) ++width;
        if (o == m) break;
    }
    return o == m;
}

gt_status read_tiff_page(const gt_file& f, const gt_file::Page& pg, uint8_t* dst) {
    const size_t bpp = f.dtype == 0 ? 2 : 1;
    const size_t row_bytes = (size_t)f.w * bpp;
    std::vector<uint8_t> buf;
    Reader R{f, {}};
    for (size_t k = 0; k < pg.off.size(); ++k) {
        const size_t r0 = k * pg.rows_per_strip;
        const size_t nr = std::min<size_t>(pg.rows_per_strip, f.h - r0);
        uint8_t* out = dst + r0 * row_bytes;
        const size_t want = nr * row_bytes;
        if (pg.compression == 1) {
            if (pg.cnt[k] < want) return set_last_error(GT_E_INVALID, "TIFF: strip too short");
            if (!R.rd(out, want, pg.off[k])) return set_last_error(GT_E_INVALID, "TIFF: " + R.err);
        } else {
            buf.resize((size_t)pg.cnt[k]);
            if (!R.rd(buf.data(), buf.size(), pg.off[k])) return set_last_error(GT_E_INVALID, "TIFF: " + R.err);
            bool ok = false;
            if (pg.compression == 32773) {
                ok = unpackbits(buf.data(), buf.size(), out, want);
            } else if (pg.compression == 5) {
                ok = unlzw(buf.data(), buf.size(), out, want);
            } else {
                uLongf dl = (uLongf)want;
                ok = uncompress(out, &dl, buf.data(), (uLong)buf.size()) == Z_OK && dl == want;
            }
            if (!ok) return set_last_error(GT_E_INVALID, "TIFF: corrupt compressed strip");
        }
        if (bpp == 2 && f.big_endian)
            for (size_t i = 0; i < want; i += 2) std::swap(out[i], out[i + 1]);
        if (pg.predictor == 2) {
            for (size_t r = 0; r < nr; ++r) {
                if (bpp == 2) {
                    uint16_t* p = (uint16_t*)(out + r * row_bytes);
                    for (uint32_t x = 1; x < f.w; ++x) p[x] = (uint16_t)(p[x] + p[x - 1]);
                } else {
                    uint8_t* p = out + r * row_bytes;
                    for (uint32_t x = 1; x < f.w; ++x) p[x] = (uint8_t)(p[x] + p[x - 1]);
                }
            }
        }
    }
    return GT_OK;
}

// ---- ETS -------------------------------------------------------------------

gt_status open_ets(gt_file& f, uint32_t channel) {
    Reader R{f, {}};
    char mg[4];
    if (!R.rd(mg, 4, 0) || std::memcmp(mg, "SIS\0", 4) != 0)
        return set_last_error(GT_E_INVALID, "ETS: missing SIS header");
    uint64_t ndim, hdr, chunks, nchunks;
    if (!R.u(12, 4, ndim) || !R.u(16, 8, hdr) || !R.u(32, 8, chunks) || !R.u(40, 4, nchunks))
        return set_last_error(GT_E_INVALID, "ETS: " + R.err);
    if (!R.rd(mg, 4, hdr) || std::memcmp(mg, "ETS\0", 4) != 0)
        return set_last_error(GT_E_INVALID, "ETS: missing ETS header");
    uint64_t dtc, d0, d1;
    if (!R.u(hdr + 8, 4, dtc) || !R.u(hdr + 28, 4, d0) || !R.u(hdr + 32, 4, d1))
        return set_last_error(GT_E_INVALID, "ETS: " + R.err);
    if (dtc != 4)
        return set_last_error(GT_E_NOT_IMPLEMENTED, "ETS: only uint16 data (type code 4) is read");
    if (ndim < 5 || ndim > 64) return set_last_error(GT_E_INVALID, "ETS: fewer than 5 dimensions");
    if (!d0 || !d1 || d0 > (1u << 20) || d1 > (1u << 20)) return set_last_error(GT_E_INVALID, "ETS: bad frame size");
    f.w = (uint32_t)d0;
    f.h = (uint32_t)d1;
    f.dtype = 0;
    std::map<uint64_t, std::vector<int64_t>> offs;   // channel -> frame offsets (-1: none)
    const uint64_t rec = 4 + 4 * ndim + 8 + 8;
    if (nchunks > (1ull << 28) || nchunks * rec > f.file_size)
        return set_last_error(GT_E_INVALID, "ETS: chunk table beyond the end of the file");
    std::vector<uint8_t> tab((size_t)(nchunks * rec));
    if (!R.rd(tab.data(), tab.size(), chunks))
        return set_last_error(GT_E_INVALID, "ETS: chunk table beyond the end of the file");
    for (uint64_t c = 0; c < nchunks; ++c) {
        const uint8_t* p = tab.data() + c * rec + 4;
        uint32_t coord[64];
        for (uint64_t k = 0; k < ndim; ++k) std::memcpy(&coord[k], p + 4 * k, 4);
        uint64_t off;
        std::memcpy(&off, p + 4 * ndim, 8);
        auto& e = offs[coord[4]];
        const uint64_t fr = coord[2];
        if (fr > (1u << 26)) return set_last_error(GT_E_INVALID, "ETS: implausible frame coordinate");
        if (fr + 1 == e.size()) e.push_back((int64_t)off);
        else if (fr + 1 > e.size()) {
            e.resize((size_t)fr + 1, -1);
            e[(size_t)fr] = (int64_t)off;
        } else {
            e[(size_t)fr] = (int64_t)off;
        }
    }
    for (const auto& kv : offs) f.channels.push_back((uint32_t)kv.first);
    if (channel == UINT32_MAX) return GT_OK;   // channel list only
    auto it = offs.find(channel);
    if (it == offs.end()) return set_last_error(GT_E_INVALID, "ETS: no such channel");
    const uint64_t fb = d0 * d1 * 2;
    for (int64_t o : it->second)
        if (o >= 0) {
            if ((uint64_t)o > f.file_size || fb > f.file_size - (uint64_t)o)
                return set_last_error(GT_E_INVALID, "ETS: frame beyond the end of the file");
            f.ets_off.push_back((uint64_t)o);
        }
    return GT_OK;
}

}  // namespace

// C entry points turn C++ exceptions into status codes.
template <class F>
static gt_status guarded_io(F&& f) {
    try {
        return f();
    } catch (const std::bad_alloc&) {
        return set_last_error(GT_E_OOM, "host allocation failed");
    } catch (...) {
        return set_last_error(GT_E_INTERNAL, "exception in the movie reader");
    }
}

extern "C" {

gt_status gt_file_open(const char* path, uint32_t channel, gt_file** out) {
    return guarded_io([&]() -> gt_status {
    if (!path || !out) return set_last_error(GT_E_INVALID, "NULL argument");
    *out = nullptr;
    std::string p(path);
    const size_t dot = p.find_last_of('.');
    std::string ext = dot == std::string::npos ? "" : p.substr(dot + 1);
    for (auto& c : ext) c = (char)std::tolower((unsigned char)c);
    int kind;
    if (ext == "tif" || ext == "tiff") kind = 0;
    else if (ext == "ets") kind = 1;
    else return set_last_error(GT_E_INVALID, "File extension '" + ext + "' not supported");
    gt_file* f = new gt_file();
    f->kind = kind;
    f->fd = open(path, O_RDONLY | O_CLOEXEC);
    if (f->fd < 0) {
        delete f;
        return set_last_error(GT_E_INVALID, std::string("Could not open file ") + path);
    }
    const off_t end = lseek(f->fd, 0, SEEK_END);
    f->file_size = end > 0 ? (uint64_t)end : 0;
    const gt_status s = kind == 0 ? open_tiff(*f) : open_ets(*f, channel);
    if (s != GT_OK) {
        delete f;
        return s;
    }
    *out = f;
    return GT_OK;
    });
}

gt_status gt_file_info(const gt_file* f, int* kind, uint64_t* n_frames, uint32_t* h, uint32_t* w,
                       int* dtype) {
    if (!f) return set_last_error(GT_E_INVALID, "NULL argument");
    if (kind) *kind = f->kind;
    if (n_frames) *n_frames = f->kind == 0 ? f->pages.size() : f->ets_off.size();
    if (h) *h = f->h;
    if (w) *w = f->w;
    if (dtype) *dtype = f->dtype;
    return GT_OK;
}

gt_status gt_file_channels(const gt_file* f, uint32_t* channels, uint32_t cap, uint32_t* n) {
    if (!f || !n) return set_last_error(GT_E_INVALID, "NULL argument");
    *n = (uint32_t)f->channels.size();
    for (uint32_t i = 0; i < *n && i < cap && channels; ++i) channels[i] = f->channels[i];
    return *n > cap && channels ? set_last_error(GT_E_INVALID, "capacity too small") : GT_OK;
}

gt_status gt_file_read(const gt_file* f, uint64_t first, uint64_t count, void* dst) {
    return guarded_io([&]() -> gt_status {
    if (!f || (count && !dst)) return set_last_error(GT_E_INVALID, "NULL argument");
    const uint64_t n = f->kind == 0 ? f->pages.size() : f->ets_off.size();
    if (first > n || count > n - first) return set_last_error(GT_E_INVALID, "frame range out of bounds");
    const size_t fb = (size_t)f->h * f->w * (f->dtype == 0 ? 2 : 1);
    for (uint64_t i = 0; i < count; ++i) {
        uint8_t* d = (uint8_t*)dst + i * fb;
        if (f->kind == 0) {
            const gt_status s = read_tiff_page(*f, f->pages[(size_t)(first + i)], d);
            if (s != GT_OK) return s;
        } else {
            Reader R{*f, {}};
            if (!R.rd(d, fb, f->ets_off[(size_t)(first + i)]))
                return set_last_error(GT_E_INVALID, "ETS: " + R.err);
        }
    }
    return GT_OK;
    });
}

void gt_file_close(gt_file* f) { delete f; }

}  // extern "C"
