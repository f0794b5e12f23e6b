// File formats of the reconstruction callers, read and written on the host:
// multi-page TIFF stacks of 32-bit float samples and numeric HDF5 datasets,
// behind the C ABI (include/laminocam.h).
//
//   include/tiff.h:31-117     tomocam::tiff::read / write (libtiff there)
//   include/hdfread.h:34-96   tomocam::h5::read_images / read_angles (libhdf5 there)
//
// Both codecs follow the published format specifications (TIFF 6.0 with the
// BigTIFF extension; the HDF5 File Format Specification, version 3), with
// zlib for the deflate streams.  What they accept:
//   TIFF  read: classic and BigTIFF, either byte order, one 32-bit IEEE float
//         sample per pixel, strips (not tiles), compression none, LZW,
//         deflate or PackBits, no predictor; every page of one size.
//         write: little-endian, one uncompressed strip per page, the tags the
//         reference's writer sets; BigTIFF when the file passes 4 GiB.
//   HDF5  superblock versions 0-3; object headers version 1 and 2; groups as
//         symbol tables or compact links; integer and IEEE float datasets of
//         either byte order (2-, 4- and 8-byte floats, 1- to 8-byte integers),
//         contiguous, compact or chunked storage (version-1 B-tree index, and
//         the single-chunk, implicit and unpaged fixed-array indexes of layout
//         version 4), filters deflate, shuffle and Fletcher-32.  Anything else
//         is refused with LC_EUNSUP and a message that names it.
#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <set>
#include <string>
#include <vector>

#include "laminocam.h"
#include "lc_common.hpp"

using lc::Error;

namespace {

// ---------------------------------------------------------------------------
// file access
// ---------------------------------------------------------------------------
class InFile {
  public:
    explicit InFile(const std::string &path) : path_(path) {
        f_ = std::fopen(path.c_str(), "rb");
        if (!f_) throw Error(LC_EINVAL, "cannot open file: " + path);
        std::fseek(f_, 0, SEEK_END);
        size_ = (uint64_t)ftello(f_);
    }
    ~InFile() {
        if (f_) std::fclose(f_);
    }
    InFile(const InFile &) = delete;
    InFile &operator=(const InFile &) = delete;
    uint64_t size() const { return size_; }
    const std::string &path() const { return path_; }
    void read(uint64_t off, void *dst, uint64_t n) const {
        if (off > size_ || n > size_ - off)
            throw Error(LC_EINVAL, path_ + ": truncated or corrupt (read past the end of the file)");
        if (fseeko(f_, (off_t)off, SEEK_SET) != 0 || std::fread(dst, 1, n, f_) != n)
            throw Error(LC_EINVAL, path_ + ": read failed");
    }
    std::vector<uint8_t> bytes(uint64_t off, uint64_t n) const {
        std::vector<uint8_t> v(n);
        if (n) read(off, v.data(), n);
        return v;
    }

  private:
    std::string path_;
    FILE *f_ = nullptr;
    uint64_t size_ = 0;
};

uint64_t le(const uint8_t *p, int n) {
    uint64_t v = 0;
    for (int i = n - 1; i >= 0; --i) v = (v << 8) | p[i];
    return v;
}
uint64_t be(const uint8_t *p, int n) {
    uint64_t v = 0;
    for (int i = 0; i < n; ++i) v = (v << 8) | p[i];
    return v;
}

// zlib stream (deflate) into exactly `want` bytes
std::vector<uint8_t> inflate_exact(const uint8_t *src, size_t n, size_t want, const std::string &what) {
    std::vector<uint8_t> out(want);
    z_stream zs{};
    if (inflateInit(&zs) != Z_OK) throw Error(LC_EINTERNAL, "zlib: inflateInit failed");
    zs.next_in = const_cast<Bytef *>(src);
    zs.avail_in = (uInt)n;
    zs.next_out = out.data();
    zs.avail_out = (uInt)want;
    const int rc = inflate(&zs, Z_FINISH);
    const size_t got = want - zs.avail_out;
    inflateEnd(&zs);
    if ((rc != Z_STREAM_END && rc != Z_OK && rc != Z_BUF_ERROR) || got != want)
        throw Error(LC_EINVAL, what + ": deflate stream does not decode to the expected size");
    return out;
}

// ---------------------------------------------------------------------------
// TIFF
// ---------------------------------------------------------------------------
enum : int {
    T_WIDTH = 256, T_LENGTH = 257, T_BPS = 258, T_COMPRESSION = 259, T_STRIP_OFFSETS = 273,
    T_SPP = 277, T_ROWS_PER_STRIP = 278, T_STRIP_BYTES = 279, T_PLANAR = 284, T_PREDICTOR = 317,
    T_TILE_WIDTH = 322, T_SAMPLE_FORMAT = 339
};

struct TiffPage {
    uint64_t width = 0, height = 0, bps = 1, spp = 1, compression = 1, planar = 1, predictor = 1,
             format = 1, rows_per_strip = UINT32_MAX;
    bool tiled = false;
    std::vector<uint64_t> offsets, counts;
};

class TiffReader {
  public:
    explicit TiffReader(const std::string &path) : f_(path) {
        uint8_t h[16] = {};
        if (f_.size() < 8) throw Error(LC_EINVAL, path + ": not a TIFF file");
        f_.read(0, h, std::min<uint64_t>(16, f_.size()));
        if (h[0] == 'I' && h[1] == 'I') big_endian_ = false;
        else if (h[0] == 'M' && h[1] == 'M') big_endian_ = true;
        else throw Error(LC_EINVAL, path + ": not a TIFF file");
        const uint64_t magic = u(h + 2, 2);
        uint64_t ifd;
        if (magic == 42) {
            ifd = u(h + 4, 4);
        } else if (magic == 43) {
            if (f_.size() < 16 || u(h + 4, 2) != 8) throw Error(LC_EINVAL, path + ": malformed BigTIFF header");
            bigtiff_ = true;
            ifd = u(h + 8, 8);
        } else {
            throw Error(LC_EINVAL, path + ": not a TIFF file");
        }
        std::set<uint64_t> seen;
        while (ifd != 0) {
            if (!seen.insert(ifd).second) throw Error(LC_EINVAL, path + ": the directory chain loops");
            if (pages_.size() >= (1u << 24)) throw Error(LC_EINVAL, path + ": too many pages");
            ifd = read_ifd(ifd);
        }
        if (pages_.empty()) throw Error(LC_EINVAL, path + ": no image directory");
        const TiffPage &p0 = pages_[0];
        for (size_t i = 0; i < pages_.size(); ++i) check_page(pages_[i], i, p0);
    }
    size_t pages() const { return pages_.size(); }
    uint64_t width() const { return pages_[0].width; }
    uint64_t height() const { return pages_[0].height; }

    // the top-left rows x cols window of page i into out (row-major floats)
    void read_page(size_t i, uint64_t rows, uint64_t cols, float *out) const {
        const TiffPage &p = pages_[i];
        const uint64_t row_bytes = p.width * 4;
        uint64_t row = 0;
        std::vector<uint8_t> raw;
        for (size_t s = 0; s < p.offsets.size() && row < rows; ++s) {
            const uint64_t nrows = std::min<uint64_t>(p.rows_per_strip, p.height - row);
            const uint64_t want = nrows * row_bytes;
            const uint64_t take = std::min(nrows, rows - row);
            if (p.compression == 1) {
                if (p.counts[s] < want) throw Error(LC_EINVAL, f_.path() + ": strip shorter than its rows");
                // uncompressed: read the needed rows of the strip straight into the output
                for (uint64_t r = 0; r < take; ++r) {
                    float *dst = out + (row + r) * cols;
                    f_.read(p.offsets[s] + r * row_bytes, dst, cols * 4);
                    if (big_endian_) swap4(dst, cols);
                }
            } else {
                raw = f_.bytes(p.offsets[s], p.counts[s]);
                const std::vector<uint8_t> dec = decode(p.compression, raw, want);
                for (uint64_t r = 0; r < take; ++r) {
                    float *dst = out + (row + r) * cols;
                    std::memcpy(dst, dec.data() + r * row_bytes, cols * 4);
                    if (big_endian_) swap4(dst, cols);
                }
            }
            row += nrows;
        }
        if (row < rows) throw Error(LC_EINVAL, f_.path() + ": strips hold fewer rows than the image length");
    }

  private:
    InFile f_;
    bool big_endian_ = false, bigtiff_ = false;
    std::vector<TiffPage> pages_;

    uint64_t u(const uint8_t *p, int n) const { return big_endian_ ? be(p, n) : le(p, n); }
    static void swap4(float *v, uint64_t n) {
        uint32_t *w = reinterpret_cast<uint32_t *>(v);
        for (uint64_t i = 0; i < n; ++i) w[i] = __builtin_bswap32(w[i]);
    }
    static int type_size(uint64_t t) {
        switch (t) {
        case 1: case 2: case 6: case 7: return 1;
        case 3: case 8: return 2;
        case 4: case 9: case 11: return 4;
        case 5: case 10: case 12: case 16: case 17: case 18: return 8;
        default: return 0;
        }
    }

    // integer values of one directory entry
    std::vector<uint64_t> values(const uint8_t *e) const {
        const uint64_t type = u(e + 2, 2);
        const uint64_t count = bigtiff_ ? u(e + 4, 8) : u(e + 4, 4);
        const int ts = type_size(type);
        if (ts == 0 || !(type == 1 || type == 3 || type == 4 || type == 16))
            throw Error(LC_EINVAL, f_.path() + ": unexpected field type " + std::to_string(type) +
                                       " for tag " + std::to_string(u(e, 2)));
        if (count > (1u << 28)) throw Error(LC_EINVAL, f_.path() + ": field count out of range");
        const uint64_t nbytes = count * (uint64_t)ts;
        const int inline_cap = bigtiff_ ? 8 : 4;
        const uint8_t *vp = e + (bigtiff_ ? 12 : 8);
        std::vector<uint8_t> ext;
        if (nbytes > (uint64_t)inline_cap) {
            ext = f_.bytes(bigtiff_ ? u(vp, 8) : u(vp, 4), nbytes);
            vp = ext.data();
        }
        std::vector<uint64_t> v(count);
        for (uint64_t i = 0; i < count; ++i) v[i] = u(vp + i * ts, ts);
        return v;
    }

    uint64_t read_ifd(uint64_t off) {
        const int cnt_bytes = bigtiff_ ? 8 : 2, ent = bigtiff_ ? 20 : 12, next_bytes = bigtiff_ ? 8 : 4;
        uint8_t cb[8];
        f_.read(off, cb, cnt_bytes);
        const uint64_t n = u(cb, cnt_bytes);
        if (n == 0 || n > 4096) throw Error(LC_EINVAL, f_.path() + ": malformed image directory");
        const std::vector<uint8_t> d = f_.bytes(off + cnt_bytes, n * ent + next_bytes);
        TiffPage p;
        for (uint64_t k = 0; k < n; ++k) {
            const uint8_t *e = d.data() + k * ent;
            const int tag = (int)u(e, 2);
            switch (tag) {
            case T_WIDTH: p.width = values(e).at(0); break;
            case T_LENGTH: p.height = values(e).at(0); break;
            case T_BPS: p.bps = values(e).at(0); break;
            case T_COMPRESSION: p.compression = values(e).at(0); break;
            case T_SPP: p.spp = values(e).at(0); break;
            case T_ROWS_PER_STRIP: p.rows_per_strip = values(e).at(0); break;
            case T_PLANAR: p.planar = values(e).at(0); break;
            case T_PREDICTOR: p.predictor = values(e).at(0); break;
            case T_SAMPLE_FORMAT: p.format = values(e).at(0); break;
            case T_STRIP_OFFSETS: p.offsets = values(e); break;
            case T_STRIP_BYTES: p.counts = values(e); break;
            case T_TILE_WIDTH: p.tiled = true; break;
            default: break;
            }
        }
        pages_.push_back(std::move(p));
        return u(d.data() + n * ent, next_bytes);
    }

    void check_page(const TiffPage &p, size_t i, const TiffPage &p0) const {
        const std::string at = f_.path() + " page " + std::to_string(i) + ": ";
        if (p.tiled) throw Error(LC_EUNSUP, at + "tiled images are not supported (strips only)");
        if (p.spp != 1 || p.bps != 32 || p.format != 3)
            throw Error(LC_EUNSUP, at + "samples must be single 32-bit IEEE floats (got " + std::to_string(p.spp) +
                                       " x " + std::to_string(p.bps) + "-bit, sample format " +
                                       std::to_string(p.format) + ")");
        if (p.compression != 1 && p.compression != 5 && p.compression != 8 && p.compression != 32946 &&
            p.compression != 32773)
            throw Error(LC_EUNSUP, at + "compression " + std::to_string(p.compression) +
                                       " is not supported (none, LZW, deflate, PackBits)");
        if (p.predictor != 1) throw Error(LC_EUNSUP, at + "predictors are not supported");
        if (p.width == 0 || p.height == 0) throw Error(LC_EINVAL, at + "empty image");
        if (p.width != p0.width || p.height != p0.height)
            throw Error(LC_EINVAL, at + "pages differ in size");
        if (p.rows_per_strip == 0) throw Error(LC_EINVAL, at + "RowsPerStrip is zero");
        const uint64_t rps = std::min<uint64_t>(p.rows_per_strip, p.height);
        const uint64_t nstrips = (p.height + rps - 1) / rps;
        if (p.offsets.size() < nstrips || p.counts.size() < p.offsets.size())
            throw Error(LC_EINVAL, at + "strip tables do not cover the image");
    }

    std::vector<uint8_t> decode(uint64_t comp, const std::vector<uint8_t> &src, uint64_t want) const {
        if (comp == 8 || comp == 32946) return inflate_exact(src.data(), src.size(), want, f_.path());
        if (comp == 5) return lzw(src, want);
        return packbits(src, want);
    }

    // TIFF LZW: MSB-first codes of 9 to 12 bits, Clear 256, EndOfInformation
    // 257, code width growing one entry early (TIFF 6.0 section 13)
    std::vector<uint8_t> lzw(const std::vector<uint8_t> &src, uint64_t want) const {
        std::vector<uint8_t> out;
        out.reserve(want);
        std::vector<uint32_t> prefix(4096), first(4096), length(4096);
        std::vector<uint8_t> last(4096);
        for (int i = 0; i < 256; ++i) {
            prefix[i] = UINT32_MAX;
            first[i] = last[i] = (uint8_t)i;
            length[i] = 1;
        }
        uint64_t bitpos = 0;
        const uint64_t nbits_total = (uint64_t)src.size() * 8;
        int width = 9;
        uint32_t next = 258;
        int64_t old = -1;
        std::vector<uint8_t> tmp;
        auto emit = [&](uint32_t code) {
            tmp.resize(length[code]);
            for (int64_t k = (int64_t)length[code] - 1; k >= 0; --k) {
                tmp[(size_t)k] = last[code];
                code = prefix[code];
            }
            out.insert(out.end(), tmp.begin(), tmp.end());
        };
        while (out.size() < want) {
            if (bitpos + (uint64_t)width > nbits_total) break;
            uint32_t code = 0;
            for (int b = 0; b < width; ++b, ++bitpos)
                code = (code << 1) | ((src[bitpos >> 3] >> (7 - (bitpos & 7))) & 1u);
            if (code == 257) break;
            if (code == 256) {
                width = 9;
                next = 258;
                old = -1;
                continue;
            }
            if (old < 0) {
                if (code > 255) throw Error(LC_EINVAL, f_.path() + ": corrupt LZW strip");
                emit(code);
                old = code;
                continue;
            }
            if (code > next || next >= 4096) throw Error(LC_EINVAL, f_.path() + ": corrupt LZW strip");
            const uint8_t head = code < next ? first[code] : first[(uint32_t)old];
            // new entry: the previous string followed by the first byte of this one
            prefix[next] = (uint32_t)old;
            last[next] = head;
            first[next] = first[(uint32_t)old];
            length[next] = length[(uint32_t)old] + 1;
            ++next;
            emit(code);
            old = code;
            if (next == 511) width = 10;
            else if (next == 1023) width = 11;
            else if (next == 2047) width = 12;
        }
        if (out.size() < want) throw Error(LC_EINVAL, f_.path() + ": LZW strip shorter than its rows");
        out.resize(want);
        return out;
    }

    std::vector<uint8_t> packbits(const std::vector<uint8_t> &src, uint64_t want) const {
        std::vector<uint8_t> out;
        out.reserve(want);
        size_t i = 0;
        while (i < src.size() && out.size() < want) {
            const int8_t n = (int8_t)src[i++];
            if (n >= 0) {
                const size_t len = (size_t)n + 1;
                if (i + len > src.size()) break;
                out.insert(out.end(), src.begin() + (long)i, src.begin() + (long)(i + len));
                i += len;
            } else if (n != -128) {
                if (i >= src.size()) break;
                out.insert(out.end(), (size_t)(1 - n), src[i++]);
            }
        }
        if (out.size() < want) throw Error(LC_EINVAL, f_.path() + ": PackBits strip shorter than its rows");
        out.resize(want);
        return out;
    }
};

class OutFile {
  public:
    explicit OutFile(const std::string &path) : path_(path) {
        f_ = std::fopen(path.c_str(), "wb");
        if (!f_) throw Error(LC_EINVAL, "cannot open file for writing: " + path);
    }
    ~OutFile() {
        if (f_) std::fclose(f_);
    }
    void write(const void *p, size_t n) {
        if (n && std::fwrite(p, 1, n, f_) != n) throw Error(LC_EINTERNAL, path_ + ": write failed");
    }
    void close() {
        FILE *f = f_;
        f_ = nullptr;
        if (std::fclose(f) != 0) throw Error(LC_EINTERNAL, path_ + ": write failed");
    }

  private:
    std::string path_;
    FILE *f_ = nullptr;
};

// little-endian TIFF: per page the samples (one strip), then the directory
void tiff_write(const std::string &path, const float *data, uint64_t pages, uint64_t rows, uint64_t cols) {
    const uint64_t page_bytes = rows * cols * 4;
    const int ntags = 12;
    const uint64_t classic_total = 8 + pages * (page_bytes + 2 + 12 * ntags + 4);
    const bool big = classic_total >= (1ull << 32);
    const uint64_t ifd_bytes = big ? 8 + 20 * ntags + 8 : 2 + 12 * ntags + 4;
    OutFile f(path);
    std::vector<uint8_t> hdr;
    auto put = [](std::vector<uint8_t> &v, uint64_t x, int n) {
        for (int i = 0; i < n; ++i) v.push_back((uint8_t)(x >> (8 * i)));
    };
    hdr.push_back('I');
    hdr.push_back('I');
    if (big) {
        put(hdr, 43, 2);
        put(hdr, 8, 2);
        put(hdr, 0, 2);
        put(hdr, 16 + page_bytes, 8);   // first directory after the first page's samples
    } else {
        put(hdr, 42, 2);
        put(hdr, 8 + page_bytes, 4);
    }
    f.write(hdr.data(), hdr.size());
    uint64_t pos = hdr.size();
    for (uint64_t i = 0; i < pages; ++i) {
        const uint64_t data_at = pos;
        f.write(data + i * rows * cols, page_bytes);   // host floats are little-endian
        pos += page_bytes;
        const uint64_t next = i + 1 < pages ? pos + ifd_bytes + page_bytes : 0;
        std::vector<uint8_t> d;
        put(d, ntags, big ? 8 : 2);
        // tag, type, value: types SHORT 3, LONG 4, LONG8 16, ascending tags
        const uint64_t long_t = big ? 16 : 4;
        const uint64_t tags[ntags][3] = {
            {T_WIDTH, 4, cols}, {T_LENGTH, 4, rows}, {T_BPS, 3, 32}, {T_COMPRESSION, 3, 1},
            {262, 3, 1},   // PhotometricInterpretation: BlackIsZero
            {T_STRIP_OFFSETS, long_t, data_at}, {274, 3, 1},   // Orientation: top-left
            {T_SPP, 3, 1}, {T_ROWS_PER_STRIP, 4, rows}, {T_STRIP_BYTES, long_t, page_bytes},
            {T_PLANAR, 3, 1}, {T_SAMPLE_FORMAT, 3, 3}};
        for (const auto &t : tags) {
            put(d, t[0], 2);
            put(d, t[1], 2);
            put(d, 1, big ? 8 : 4);
            const int vb = big ? 8 : 4;
            const int tsz = t[1] == 3 ? 2 : t[1] == 4 ? 4 : 8;
            put(d, t[2], tsz);
            put(d, 0, vb - tsz);
        }
        put(d, next, big ? 8 : 4);
        f.write(d.data(), d.size());
        pos += d.size();
    }
    f.close();
}

// ---------------------------------------------------------------------------
// HDF5
// ---------------------------------------------------------------------------
constexpr uint64_t H5_UNDEF = UINT64_MAX;

struct H5Msg {
    int type;
    int flags;
    std::vector<uint8_t> data;
};

struct H5Type {
    int cls = -1;   // 0 integer, 1 float
    int size = 0;
    bool big_endian = false, is_signed = false;
};

struct H5Filter {
    int id;
    std::vector<uint32_t> cd;
};

struct H5Dataset {
    std::vector<uint64_t> dims;
    H5Type type;
    int layout = -1;   // 0 compact, 1 contiguous, 2 chunked
    uint64_t addr = H5_UNDEF, size = 0;
    std::vector<uint8_t> compact;
    int chunk_index = 0;   // 0 version-1 B-tree; 1 single, 2 implicit, 3 fixed array
    std::vector<uint64_t> chunk;   // chunk dims (rank entries)
    uint64_t single_size = 0;
    uint32_t single_mask = 0;
    bool single_filtered = false;
    std::vector<H5Filter> filters;
    std::vector<uint8_t> fill;
};

class H5Reader {
  public:
    explicit H5Reader(const std::string &path) : f_(path) {
        // the superblock is at 0, 512, 1024, ... (a user block may precede it)
        static const uint8_t sig[8] = {0x89, 'H', 'D', 'F', '\r', '\n', 0x1a, '\n'};
        uint64_t at = 0;
        for (;; at = at ? at * 2 : 512) {
            if (at + 8 > f_.size()) throw Error(LC_EINVAL, path + ": not an HDF5 file");
            uint8_t s[8];
            f_.read(at, s, 8);
            if (!std::memcmp(s, sig, 8)) break;
        }
        const std::vector<uint8_t> h = f_.bytes(at, std::min<uint64_t>(f_.size() - at, 128));
        const int ver = h.at(8);
        if (ver == 0 || ver == 1) {
            so_ = h.at(13);
            sl_ = h.at(14);
            check_sizes();
            size_t p = 24 + (ver == 1 ? 4 : 0);
            base_ = rd(h, p, so_);
            p += 4 * so_;   // base, free-space, end-of-file, driver info
            // root group symbol table entry: name offset, object header address
            root_ = rd(h, p + so_, so_);
        } else if (ver == 2 || ver == 3) {
            so_ = h.at(9);
            sl_ = h.at(10);
            check_sizes();
            size_t p = 12;
            base_ = rd(h, p, so_);
            root_ = rd(h, p + 3 * so_, so_);   // after base, extension, end-of-file
        } else {
            throw Error(LC_EUNSUP, path + ": HDF5 superblock version " + std::to_string(ver) + " is not supported");
        }
        if (base_ == H5_UNDEF || base_ > at) base_ = at;
    }

    H5Dataset open(const std::string &name) {
        uint64_t oh = root_;
        size_t i = 0;
        std::string walked;
        while (i < name.size()) {
            while (i < name.size() && name[i] == '/') ++i;
            if (i >= name.size()) break;
            size_t j = name.find('/', i);
            if (j == std::string::npos) j = name.size();
            const std::string part = name.substr(i, j - i);
            walked += "/" + part;
            oh = lookup(oh, part, walked);
            i = j;
        }
        return dataset(oh, name);
    }

    // the dataset's elements, converted, into out (count = product of dims)
    template <typename D> void read(const H5Dataset &ds, D *out) {
        uint64_t n = 1;
        for (uint64_t d : ds.dims) n *= d;
        const int es = ds.type.size;
        if (ds.layout == 0) {
            if (ds.compact.size() < n * es) throw Error(LC_EINVAL, f_.path() + ": compact data too short");
            convert(ds.type, ds.compact.data(), n, out);
        } else if (ds.layout == 1) {
            if (ds.addr == H5_UNDEF) {
                fill_out(ds, out, n);
                return;
            }
            // stream in blocks so a large dataset needs no second full copy
            const uint64_t block = (64u << 20) / es;
            std::vector<uint8_t> buf;
            for (uint64_t k = 0; k < n; k += block) {
                const uint64_t m = std::min(block, n - k);
                buf.resize(m * es);
                f_.read(base_ + ds.addr + k * es, buf.data(), m * es);
                convert(ds.type, buf.data(), m, out + k);
            }
        } else {
            read_chunked(ds, out, n);
        }
    }

  private:
    InFile f_;
    int so_ = 8, sl_ = 8;
    uint64_t base_ = 0, root_ = H5_UNDEF;

    void check_sizes() const {
        if ((so_ != 2 && so_ != 4 && so_ != 8) || (sl_ != 2 && sl_ != 4 && sl_ != 8))
            throw Error(LC_EINVAL, f_.path() + ": bad offset/length sizes in the superblock");
    }
    static uint64_t rd(const std::vector<uint8_t> &v, size_t p, int n) {
        if (p + (size_t)n > v.size()) throw Error(LC_EINVAL, "HDF5: structure truncated");
        const uint64_t x = le(v.data() + p, n);
        if (n < 8 && x == (1ull << (8 * n)) - 1) return H5_UNDEF;
        return x;
    }
    uint64_t fa(uint64_t addr) const {   // file offset of an address
        if (addr == H5_UNDEF) throw Error(LC_EINVAL, f_.path() + ": undefined address");
        return base_ + addr;
    }

    // messages of the object header at addr, continuation blocks followed
    std::vector<H5Msg> header(uint64_t addr) {
        std::vector<H5Msg> out;
        uint8_t sig[4];
        f_.read(fa(addr), sig, 4);
        if (!std::memcmp(sig, "OHDR", 4)) {
            const std::vector<uint8_t> h = f_.bytes(fa(addr), 16);
            const int flags = h[5];
            size_t p = 6;
            if (flags & 0x20) p += 16;
            if (flags & 0x10) p += 4;
            const int szb = 1 << (flags & 3);
            const uint64_t chunk0 = rd(f_.bytes(fa(addr) + p, szb), 0, szb);
            p += szb;
            std::vector<std::pair<uint64_t, uint64_t>> blocks{{fa(addr) + p, chunk0}};
            const bool order = flags & 0x04;
            for (size_t b = 0; b < blocks.size(); ++b) {
                const std::vector<uint8_t> d = f_.bytes(blocks[b].first, blocks[b].second);
                size_t q = b == 0 ? 0 : 4;   // continuation blocks start with OCHK
                const size_t end = d.size() - (b == 0 ? 0 : 4);   // and end with a checksum
                const size_t mh = order ? 6 : 4;
                while (q + mh <= end) {
                    const int type = d[q];
                    const size_t sz = (size_t)le(&d[q + 1], 2);
                    const int mf = d[q + 3];
                    q += mh;
                    if (q + sz > end) break;
                    H5Msg m{type, mf, std::vector<uint8_t>(d.begin() + (long)q, d.begin() + (long)(q + sz))};
                    q += sz;
                    if (type == 0x10) {
                        blocks.push_back({fa(rd(m.data, 0, so_)), rd(m.data, so_, sl_)});
                        if (blocks.size() > 4096) throw Error(LC_EINVAL, f_.path() + ": header continuation loop");
                    } else {
                        out.push_back(std::move(m));
                    }
                }
            }
            return out;
        }
        const std::vector<uint8_t> h = f_.bytes(fa(addr), 16);
        if (h[0] != 1) throw Error(LC_EUNSUP, f_.path() + ": object header version " + std::to_string(h[0]));
        const uint64_t nmsg = le(&h[2], 2);
        std::vector<std::pair<uint64_t, uint64_t>> blocks{{fa(addr) + 16, le(&h[8], 4)}};
        for (size_t b = 0; b < blocks.size() && out.size() < nmsg; ++b) {
            const std::vector<uint8_t> d = f_.bytes(blocks[b].first, blocks[b].second);
            size_t q = 0;
            while (q + 8 <= d.size()) {
                const int type = (int)le(&d[q], 2);
                const size_t sz = (size_t)le(&d[q + 2], 2);
                const int mf = d[q + 4];
                q += 8;
                if (q + sz > d.size()) break;
                H5Msg m{type, mf, std::vector<uint8_t>(d.begin() + (long)q, d.begin() + (long)(q + sz))};
                q += (sz + 7) & ~size_t(7);
                if (type == 0x10) {
                    blocks.push_back({fa(rd(m.data, 0, so_)), rd(m.data, so_, sl_)});
                    if (blocks.size() > 4096) throw Error(LC_EINVAL, f_.path() + ": header continuation loop");
                } else {
                    out.push_back(std::move(m));
                }
            }
        }
        return out;
    }

    // object header address of the link `name` in the group at oh
    uint64_t lookup(uint64_t oh, const std::string &name, const std::string &walked) {
        const std::vector<H5Msg> msgs = header(oh);
        for (const H5Msg &m : msgs) {
            if (m.type == 0x11) {   // symbol table: v1 B-tree of symbol nodes + local heap
                const uint64_t bt = rd(m.data, 0, so_), heap = rd(m.data, so_, so_);
                const std::vector<uint8_t> hh = f_.bytes(fa(heap), 8 + 2 * sl_ + so_);
                if (std::memcmp(hh.data(), "HEAP", 4)) throw Error(LC_EINVAL, f_.path() + ": bad local heap");
                const uint64_t heap_size = rd(hh, 8, sl_);
                const uint64_t heap_data = rd(hh, 8 + 2 * sl_, so_);
                const std::vector<uint8_t> names = f_.bytes(fa(heap_data), heap_size);
                uint64_t found = H5_UNDEF;
                group_btree(bt, names, name, found, 0);
                if (found != H5_UNDEF) return found;
                throw Error(LC_EINVAL, f_.path() + ": no object " + walked);
            }
        }
        for (const H5Msg &m : msgs) {
            if (m.type == 0x02) {   // link info: a defined fractal heap means dense link storage
                const int fl = m.data.at(1);
                const size_t p = 2 + ((fl & 1) ? 8 : 0);
                if (rd(m.data, p, so_) != H5_UNDEF)
                    throw Error(LC_EUNSUP, f_.path() + ": dense link storage (groups of many links) is not supported");
            }
        }
        for (const H5Msg &m : msgs) {
            if (m.type != 0x06) continue;
            const std::vector<uint8_t> &d = m.data;
            const int fl = d.at(1);
            size_t p = 2;
            int ltype = 0;
            if (fl & 0x08) ltype = d.at(p++);
            if (fl & 0x04) p += 8;
            if (fl & 0x10) ++p;
            const int lb = 1 << (fl & 3);
            const uint64_t len = le(&d.at(p), lb);
            p += lb;
            if (p + len > d.size()) throw Error(LC_EINVAL, f_.path() + ": bad link message");
            if (std::string(d.begin() + (long)p, d.begin() + (long)(p + len)) != name) continue;
            p += len;
            if (ltype != 0) throw Error(LC_EUNSUP, f_.path() + ": " + walked + " is a soft or external link");
            return rd(d, p, so_);
        }
        throw Error(LC_EINVAL, f_.path() + ": no object " + walked);
    }

    void group_btree(uint64_t node, const std::vector<uint8_t> &names, const std::string &name, uint64_t &found,
                     int depth) {
        if (depth > 64) throw Error(LC_EINVAL, f_.path() + ": group B-tree too deep");
        const std::vector<uint8_t> h = f_.bytes(fa(node), 8 + 2 * so_);
        if (std::memcmp(h.data(), "TREE", 4) || h[4] != 0) throw Error(LC_EINVAL, f_.path() + ": bad group B-tree");
        const int level = h[5];
        const uint64_t used = le(&h[6], 2);
        const std::vector<uint8_t> kc = f_.bytes(fa(node) + 8 + 2 * so_, used * (sl_ + so_) + sl_);
        for (uint64_t i = 0; i < used && found == H5_UNDEF; ++i) {
            const uint64_t child = rd(kc, i * (sl_ + so_) + sl_, so_);
            if (level > 0) {
                group_btree(child, names, name, found, depth + 1);
                continue;
            }
            const std::vector<uint8_t> sn = f_.bytes(fa(child), 8);
            if (std::memcmp(sn.data(), "SNOD", 4)) throw Error(LC_EINVAL, f_.path() + ": bad symbol node");
            const uint64_t nsym = le(&sn[6], 2);
            const size_t es = 2 * so_ + 24;
            const std::vector<uint8_t> ents = f_.bytes(fa(child) + 8, nsym * es);
            for (uint64_t k = 0; k < nsym; ++k) {
                const uint64_t off = rd(ents, k * es, so_);
                if (off >= names.size()) continue;
                const char *s = reinterpret_cast<const char *>(names.data() + off);
                const size_t sl = strnlen(s, names.size() - off);
                if (std::string(s, sl) == name) {
                    found = rd(ents, k * es + so_, so_);
                    return;
                }
            }
        }
    }

    H5Dataset dataset(uint64_t oh, const std::string &name) {
        const std::vector<H5Msg> msgs = header(oh);
        H5Dataset ds;
        bool has_space = false, has_type = false;
        std::vector<uint8_t> old_fill;
        for (const H5Msg &m : msgs) {
            const std::vector<uint8_t> &d = m.data;
            if ((m.flags & 0x02) && (m.type == 0x03 || m.type == 0x01))
                throw Error(LC_EUNSUP, f_.path() + ": " + name + " uses a shared (committed) datatype or dataspace");
            switch (m.type) {
            case 0x01: {   // dataspace
                const int ver = d.at(0), rank = d.at(1);
                const size_t p = ver == 1 ? 8 : 4;
                if (ver == 2 && d.at(3) == 2) throw Error(LC_EINVAL, f_.path() + ": " + name + " has a null dataspace");
                for (int k = 0; k < rank; ++k) ds.dims.push_back(rd(d, p + (size_t)k * sl_, sl_));
                has_space = true;
                break;
            }
            case 0x03: {   // datatype
                ds.type.cls = d.at(0) & 0x0f;
                const int b0 = d.at(1);
                ds.type.size = (int)le(&d.at(4), 4);
                ds.type.big_endian = b0 & 1;
                if (ds.type.cls == 0) {
                    ds.type.is_signed = b0 & 0x08;
                    if (ds.type.size != 1 && ds.type.size != 2 && ds.type.size != 4 && ds.type.size != 8)
                        ds.type.cls = -2;
                } else if (ds.type.cls == 1) {
                    const uint64_t off = le(&d.at(8), 2), prec = le(&d.at(10), 2);
                    if ((b0 & 0x40) || off != 0 || (int)prec != 8 * ds.type.size ||
                        (ds.type.size != 2 && ds.type.size != 4 && ds.type.size != 8))
                        ds.type.cls = -2;
                }
                has_type = true;
                break;
            }
            case 0x04:   // fill value (old form): size, value
                old_fill.assign(d.begin() + 4, d.begin() + 4 + (long)std::min<uint64_t>(le(&d.at(0), 4), d.size() - 4));
                break;
            case 0x05: {   // fill value
                const int ver = d.at(0);
                if (ver == 1 || ver == 2) {
                    const bool defined = d.at(3);
                    if ((ver == 1 || defined) && d.size() >= 8) {
                        const uint64_t sz = le(&d.at(4), 4);
                        if (sz && 8 + sz <= d.size()) ds.fill.assign(d.begin() + 8, d.begin() + 8 + (long)sz);
                    }
                } else if (ver == 3) {
                    if ((d.at(1) & 0x20) && d.size() >= 6) {
                        const uint64_t sz = le(&d.at(2), 4);
                        if (sz && 6 + sz <= d.size()) ds.fill.assign(d.begin() + 6, d.begin() + 6 + (long)sz);
                    }
                }
                break;
            }
            case 0x08: layout(d, ds, name); break;
            case 0x0B: filters(d, ds); break;
            default: break;
            }
        }
        if (!has_space || !has_type || ds.layout < 0)
            throw Error(LC_EINVAL, f_.path() + ": " + name + " is not a dataset");
        if (ds.type.cls < 0 || ds.type.cls > 1)
            throw Error(LC_EUNSUP, f_.path() + ": " + name + " is not an integer or IEEE float dataset");
        if (ds.fill.empty()) ds.fill = old_fill;
        if (!ds.fill.empty() && (int)ds.fill.size() != ds.type.size) ds.fill.clear();
        if (ds.layout == 2 && ds.chunk.size() != ds.dims.size())
            throw Error(LC_EINVAL, f_.path() + ": " + name + " chunk rank differs from its dataspace");
        return ds;
    }

    void layout(const std::vector<uint8_t> &d, H5Dataset &ds, const std::string &name) {
        const int ver = d.at(0);
        if (ver == 1 || ver == 2) {
            const int nd = d.at(1);
            ds.layout = d.at(2);
            size_t p = 8;
            if (ds.layout != 0) {
                ds.addr = rd(d, p, so_);
                p += so_;
            }
            std::vector<uint64_t> dd;
            for (int k = 0; k < nd; ++k) dd.push_back(le(&d.at(p + 4 * k), 4));
            p += 4 * (size_t)nd;
            if (ds.layout == 2) {   // the last dimension is the element size
                ds.chunk.assign(dd.begin(), dd.end() - 1);
            } else if (ds.layout == 0) {
                const uint64_t sz = le(&d.at(p), 4);
                ds.compact.assign(d.begin() + (long)p + 4, d.begin() + (long)(p + 4 + sz));
            }
        } else if (ver == 3 || ver == 4) {
            ds.layout = d.at(1);
            size_t p = 2;
            if (ds.layout == 0) {
                const uint64_t sz = le(&d.at(p), 2);
                ds.compact.assign(d.begin() + (long)p + 2, d.begin() + (long)(p + 2 + sz));
            } else if (ds.layout == 1) {
                ds.addr = rd(d, p, so_);
                ds.size = rd(d, p + so_, sl_);
            } else if (ds.layout == 2 && ver == 3) {
                const int nd = d.at(p);
                ds.addr = rd(d, p + 1, so_);
                p += 1 + so_;
                for (int k = 0; k + 1 < nd; ++k) ds.chunk.push_back(le(&d.at(p + 4 * k), 4));
            } else if (ds.layout == 2) {
                const int fl = d.at(p), nd = d.at(p + 1), enc = d.at(p + 2);
                p += 3;
                for (int k = 0; k + 1 < nd; ++k) ds.chunk.push_back(le(&d.at(p + (size_t)enc * k), enc));
                p += (size_t)enc * nd;
                const int idx = d.at(p++);
                if (idx == 1) {
                    ds.chunk_index = 1;
                    if (fl & 0x02) {
                        ds.single_filtered = true;
                        ds.single_size = rd(d, p, sl_);
                        ds.single_mask = (uint32_t)le(&d.at(p + sl_), 4);
                        p += sl_ + 4;
                    }
                } else if (idx == 2) {
                    ds.chunk_index = 2;
                } else if (idx == 3) {
                    ds.chunk_index = 3;
                    p += 1;   // page bits
                } else {
                    throw Error(LC_EUNSUP, f_.path() + ": " + name + " uses chunk index type " + std::to_string(idx) +
                                               " (extensible array or v2 B-tree), not supported");
                }
                ds.addr = rd(d, p, so_);
            } else {
                throw Error(LC_EUNSUP, f_.path() + ": " + name + " uses virtual storage, not supported");
            }
        } else {
            throw Error(LC_EUNSUP, f_.path() + ": layout message version " + std::to_string(ver));
        }
    }

    void filters(const std::vector<uint8_t> &d, H5Dataset &ds) {
        const int ver = d.at(0), n = d.at(1);
        size_t p = ver == 1 ? 8 : 2;
        for (int k = 0; k < n; ++k) {
            H5Filter fl;
            fl.id = (int)le(&d.at(p), 2);
            p += 2;
            uint64_t namelen = 0;
            if (ver == 1 || fl.id >= 256) {
                namelen = le(&d.at(p), 2);
                p += 2;
            }
            p += 2;   // flags
            const uint64_t nv = le(&d.at(p), 2);
            p += 2;
            if (ver == 1) namelen = (namelen + 7) & ~uint64_t(7);
            p += namelen;
            for (uint64_t v = 0; v < nv; ++v) fl.cd.push_back((uint32_t)le(&d.at(p + 4 * v), 4));
            p += 4 * nv;
            if (ver == 1 && (nv & 1)) p += 4;
            ds.filters.push_back(fl);
        }
    }

    // undo the filter pipeline (in reverse order) except filters masked off
    std::vector<uint8_t> unfilter(const H5Dataset &ds, std::vector<uint8_t> buf, uint32_t mask, uint64_t want) {
        for (int k = (int)ds.filters.size() - 1; k >= 0; --k) {
            if (mask & (1u << k)) continue;
            const H5Filter &fl = ds.filters[(size_t)k];
            if (fl.id == 1) {
                buf = inflate_exact(buf.data(), buf.size(), want, f_.path());
            } else if (fl.id == 2) {
                const uint64_t es = fl.cd.empty() ? (uint64_t)ds.type.size : fl.cd[0];
                const uint64_t ne = buf.size() / es;
                std::vector<uint8_t> o(buf.size());
                for (uint64_t b = 0; b < es; ++b)
                    for (uint64_t i = 0; i < ne; ++i) o[i * es + b] = buf[b * ne + i];
                std::copy(buf.begin() + (long)(ne * es), buf.end(), o.begin() + (long)(ne * es));
                buf.swap(o);
            } else if (fl.id == 3) {
                if (buf.size() < 4) throw Error(LC_EINVAL, f_.path() + ": Fletcher-32 chunk too short");
                buf.resize(buf.size() - 4);
            } else {
                throw Error(LC_EUNSUP, f_.path() + ": HDF5 filter " + std::to_string(fl.id) +
                                           " is not supported (deflate, shuffle, Fletcher-32)");
            }
        }
        return buf;
    }

    struct ChunkRef {
        uint64_t addr, size;
        uint32_t mask;
        std::vector<uint64_t> off;   // element offsets of the chunk's first element
    };

    void chunk_btree(uint64_t node, int rank, std::vector<ChunkRef> &out, int depth) {
        if (depth > 64) throw Error(LC_EINVAL, f_.path() + ": chunk B-tree too deep");
        const std::vector<uint8_t> h = f_.bytes(fa(node), 8 + 2 * so_);
        if (std::memcmp(h.data(), "TREE", 4) || h[4] != 1) throw Error(LC_EINVAL, f_.path() + ": bad chunk B-tree");
        const int level = h[5];
        const uint64_t used = le(&h[6], 2);
        const size_t ks = 8 + 8 * (size_t)(rank + 1);
        const std::vector<uint8_t> kc = f_.bytes(fa(node) + 8 + 2 * so_, used * (ks + so_) + ks);
        for (uint64_t i = 0; i < used; ++i) {
            const size_t kp = i * (ks + so_);
            const uint64_t child = rd(kc, kp + ks, so_);
            if (level > 0) {
                chunk_btree(child, rank, out, depth + 1);
                continue;
            }
            ChunkRef c{child, le(&kc[kp], 4), (uint32_t)le(&kc[kp + 4], 4), {}};
            for (int r = 0; r < rank; ++r) c.off.push_back(le(&kc[kp + 8 + 8 * (size_t)r], 8));
            out.push_back(std::move(c));
        }
    }

    template <typename D> void fill_out(const H5Dataset &ds, D *out, uint64_t n) {
        D v = D(0);
        if (!ds.fill.empty()) convert(ds.type, ds.fill.data(), 1, &v);
        std::fill(out, out + n, v);
    }

    template <typename D> void read_chunked(const H5Dataset &ds, D *out, uint64_t n) {
        const int rank = (int)ds.dims.size();
        const int es = ds.type.size;
        uint64_t cn = 1;
        for (uint64_t c : ds.chunk) cn *= c;
        if (cn == 0) throw Error(LC_EINVAL, f_.path() + ": zero chunk size");
        const uint64_t cbytes = cn * es;
        fill_out(ds, out, n);
        if (ds.addr == H5_UNDEF || n == 0) return;
        // chunk grid (row-major) for the implicit and fixed-array indexes
        std::vector<uint64_t> grid(rank);
        uint64_t nchunks = 1;
        for (int r = 0; r < rank; ++r) {
            grid[r] = (ds.dims[r] + ds.chunk[r] - 1) / ds.chunk[r];
            nchunks *= grid[r];
        }
        auto grid_offsets = [&](uint64_t k) {
            std::vector<uint64_t> off(rank);
            for (int r = rank - 1; r >= 0; --r) {
                off[r] = (k % grid[r]) * ds.chunk[r];
                k /= grid[r];
            }
            return off;
        };
        std::vector<ChunkRef> refs;
        if (ds.chunk_index == 0) {
            chunk_btree(ds.addr, rank, refs, 0);
        } else if (ds.chunk_index == 1) {
            refs.push_back({ds.addr, ds.single_filtered ? ds.single_size : cbytes, ds.single_mask,
                            std::vector<uint64_t>(rank, 0)});
        } else if (ds.chunk_index == 2) {
            for (uint64_t k = 0; k < nchunks; ++k) refs.push_back({ds.addr + k * cbytes, cbytes, 0, grid_offsets(k)});
        } else {
            const std::vector<uint8_t> fh = f_.bytes(fa(ds.addr), 12 + sl_ + so_);
            if (std::memcmp(fh.data(), "FAHD", 4)) throw Error(LC_EINVAL, f_.path() + ": bad fixed-array header");
            const int client = fh[5], esz = fh[6], page_bits = fh[7];
            const uint64_t nent = rd(fh, 8, sl_);
            const uint64_t dblk = rd(fh, 8 + sl_, so_);
            if (nent > (1ull << page_bits))
                throw Error(LC_EUNSUP, f_.path() + ": paged fixed-array chunk index is not supported");
            const std::vector<uint8_t> db = f_.bytes(fa(dblk), 6 + so_ + nent * esz);
            if (std::memcmp(db.data(), "FADB", 4)) throw Error(LC_EINVAL, f_.path() + ": bad fixed-array data block");
            for (uint64_t k = 0; k < std::min(nent, nchunks); ++k) {
                const size_t p = 6 + so_ + k * esz;
                const uint64_t a = rd(db, p, so_);
                if (a == H5_UNDEF) continue;
                if (client == 0) {
                    refs.push_back({a, cbytes, 0, grid_offsets(k)});
                } else {
                    const int lb = esz - so_ - 4;
                    refs.push_back({a, le(&db[p + so_], lb), (uint32_t)le(&db[p + so_ + lb], 4), grid_offsets(k)});
                }
            }
        }
        std::vector<D> conv(cn);
        for (const ChunkRef &c : refs) {
            std::vector<uint8_t> raw = f_.bytes(fa(c.addr), c.size);
            raw = unfilter(ds, std::move(raw), c.mask, cbytes);
            if (raw.size() < cbytes) throw Error(LC_EINVAL, f_.path() + ": chunk shorter than its extent");
            convert(ds.type, raw.data(), cn, conv.data());
            // copy the part of the chunk inside the dataset, one innermost run at a time
            const uint64_t inner = ds.chunk[rank - 1];
            const uint64_t runs = cn / inner;
            for (uint64_t rr = 0; rr < runs; ++rr) {
                uint64_t t = rr, dst = 0, stride = 1;
                bool inside = true;
                std::vector<uint64_t> idx(rank);
                for (int r = rank - 2; r >= 0; --r) {
                    idx[r] = t % ds.chunk[r];
                    t /= ds.chunk[r];
                }
                for (int r = rank - 1; r >= 0; --r) {
                    const uint64_t g = (r == rank - 1 ? 0 : idx[r]) + c.off[r];
                    if (g >= ds.dims[r]) inside = false;
                    dst += g * stride;
                    stride *= ds.dims[r];
                }
                if (!inside) continue;
                const uint64_t len = std::min(inner, ds.dims[rank - 1] - c.off[rank - 1]);
                std::copy(conv.begin() + (long)(rr * inner), conv.begin() + (long)(rr * inner + len), out + dst);
            }
        }
    }

    static double half_to_double(uint16_t h) {
        const int s = h >> 15, e = (h >> 10) & 0x1f, m = h & 0x3ff;
        double v;
        if (e == 0) v = std::ldexp((double)m, -24);
        else if (e == 31) v = m ? NAN : INFINITY;
        else v = std::ldexp((double)(m | 0x400), e - 25);
        return s ? -v : v;
    }

    template <typename D> static void convert(const H5Type &t, const uint8_t *src, uint64_t n, D *dst) {
        const int es = t.size;
        for (uint64_t i = 0; i < n; ++i) {
            const uint8_t *p = src + i * es;
            const uint64_t bits = t.big_endian ? be(p, es) : le(p, es);
            if (t.cls == 1) {
                if (es == 4) {
                    float f;
                    const uint32_t b = (uint32_t)bits;
                    std::memcpy(&f, &b, 4);
                    dst[i] = (D)f;
                } else if (es == 8) {
                    double g;
                    std::memcpy(&g, &bits, 8);
                    dst[i] = (D)g;
                } else {
                    dst[i] = (D)half_to_double((uint16_t)bits);
                }
            } else if (t.is_signed) {
                const int sh = 64 - 8 * es;   // sign-extend the es-byte value
                dst[i] = (D)(sh ? ((int64_t)(bits << sh)) >> sh : (int64_t)bits);
            } else {
                dst[i] = (D)bits;
            }
        }
    }
};

} // namespace

extern "C" {

int lc_tiff_info(const char *path, int64_t *pages, int64_t *height, int64_t *width) {
    return lc::c_guard([&] {
        LC_REQUIRE(path && pages && height && width, "lc_tiff_info: NULL argument");
        TiffReader t(path);
        *pages = (int64_t)t.pages();
        *height = (int64_t)t.height();
        *width = (int64_t)t.width();
    });
}

int lc_tiff_read(const char *path, float *out, int64_t pages, int64_t rows, int64_t cols) {
    return lc::c_guard([&] {
        LC_REQUIRE(path && out, "lc_tiff_read: NULL argument");
        TiffReader t(path);
        LC_REQUIRE(pages >= 0 && (uint64_t)pages <= t.pages() && rows >= 1 && (uint64_t)rows <= t.height() &&
                       cols >= 1 && (uint64_t)cols <= t.width(),
                   "lc_tiff_read: the window exceeds the file's pages or image size");
        for (int64_t i = 0; i < pages; ++i) t.read_page((size_t)i, (uint64_t)rows, (uint64_t)cols, out + i * rows * cols);
    });
}

int lc_tiff_write(const char *path, const float *data, int64_t pages, int64_t rows, int64_t cols) {
    return lc::c_guard([&] {
        LC_REQUIRE(path && data, "lc_tiff_write: NULL argument");
        LC_REQUIRE(pages >= 1 && rows >= 1 && cols >= 1 && rows <= UINT32_MAX && cols <= UINT32_MAX,
                   "lc_tiff_write: pages, rows and cols must be positive (rows, cols below 2^32)");
        tiff_write(path, data, (uint64_t)pages, (uint64_t)rows, (uint64_t)cols);
    });
}

int lc_h5_shape(const char *path, const char *dataset, int32_t *rank, int64_t *dims) {
    return lc::c_guard([&] {
        LC_REQUIRE(path && dataset && rank && dims, "lc_h5_shape: NULL argument");
        H5Reader h(path);
        const H5Dataset ds = h.open(dataset);
        LC_REQUIRE(ds.dims.size() <= 32, "lc_h5_shape: rank above 32");
        *rank = (int32_t)ds.dims.size();
        for (size_t k = 0; k < ds.dims.size(); ++k) dims[k] = (int64_t)ds.dims[k];
    });
}

int lc_h5_read(const char *path, const char *dataset, void *out, int64_t count, int32_t f64) {
    return lc::c_guard([&] {
        LC_REQUIRE(path && dataset && out, "lc_h5_read: NULL argument");
        H5Reader h(path);
        const H5Dataset ds = h.open(dataset);
        uint64_t n = 1;
        for (uint64_t d : ds.dims) n *= d;
        LC_REQUIRE((uint64_t)count == n, "lc_h5_read: count differs from the dataset's element count");
        if (f64) h.read(ds, static_cast<double *>(out));
        else h.read(ds, static_cast<float *>(out));
    });
}

} // extern "C"
