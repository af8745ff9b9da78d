// kyto's store format v0 on the host side of the library: the reader of a tiled CSR
// layer (docs/format.md §4 of the reference; crates/kyto-store/src/tile_reader.rs,
// tile_decode.rs, layer_open.rs) that decodes the store's zstd-compressed, sharded
// Zarr v3 tiles straight into the pinned buffers of the block stream (csr_stream), and
// the writers and readers of cell/gene embeddings (§6, embedding_io.rs) and cell graphs
// (§7, graph_io.rs) that hold the path's outputs (X_pca, distances, connectivities).
//
// zstd is the system's libzstd (its stable one-shot API, declared below); crc32c,
// JSON, the Zarr v3 array metadata, the sharding_indexed codec and the tile layout are
// this file's own.
#include <dirent.h>
#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <array>
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "common.hpp"
#include "store.hpp"

extern "C" {
typedef struct ZSTD_DCtx_s ZSTD_DCtx;
typedef struct ZSTD_CCtx_s ZSTD_CCtx;
ZSTD_DCtx *ZSTD_createDCtx(void);
size_t ZSTD_freeDCtx(ZSTD_DCtx *dctx);
size_t ZSTD_decompressDCtx(ZSTD_DCtx *dctx, void *dst, size_t dst_capacity, const void *src, size_t src_size);
ZSTD_CCtx *ZSTD_createCCtx(void);
size_t ZSTD_freeCCtx(ZSTD_CCtx *cctx);
size_t ZSTD_compressCCtx(ZSTD_CCtx *cctx, void *dst, size_t dst_capacity, const void *src, size_t src_size,
                         int level);
size_t ZSTD_compressBound(size_t src_size);
unsigned ZSTD_isError(size_t code);
const char *ZSTD_getErrorName(size_t code);
}

namespace ky {
namespace {

constexpr uint64_t MISSING = ~uint64_t(0);     // a shard index entry of an absent inner chunk
constexpr uint64_t OFFSETS_CHUNK = 65536;     // tile_offsets and graph indptr chunks (§4.4, §7)
constexpr uint64_t MAX_BUFFERED = 1ull << 31;  // bytes of one chunk or shard (§4.5)

// ---------------------------------------------------------------- utilities --

// Tasks on the library's host threads; the first failure is rethrown on the caller.
template <typename F>
void parallel_tasks(size_t n, F &&f) {
    std::mutex m;
    std::exception_ptr first;
    host_parallel(n, [&](size_t t) {
        try {
            f(t);
        } catch (...) {
            std::lock_guard<std::mutex> g(m);
            if (!first) first = std::current_exception();
        }
    });
    if (first) std::rethrow_exception(first);
}

[[noreturn]] void io_error(const std::string &operation, const std::string &path, const std::string &detail) {
    throw Error(KY_E110, "could not " + operation + " '" + path + "': " + detail);
}

// CRC-32C (Castagnoli, reflected polynomial 0x82F63B78), the sharding index checksum.
uint32_t crc32c(const uint8_t *p, size_t n) {
    static const std::array<uint32_t, 256> table = [] {
        std::array<uint32_t, 256> t{};
        for (uint32_t i = 0; i < 256; i++) {
            uint32_t c = i;
            for (int k = 0; k < 8; k++) c = (c >> 1) ^ (0x82F63B78u & (0u - (c & 1u)));
            t[i] = c;
        }
        return t;
    }();
    uint32_t c = 0xFFFFFFFFu;
    for (size_t i = 0; i < n; i++) c = table[(c ^ p[i]) & 0xFFu] ^ (c >> 8);
    return c ^ 0xFFFFFFFFu;
}

// One zstd decompression / compression context per host thread, kept for its life.
struct ZstdContexts {
    ZSTD_DCtx *d = nullptr;
    ZSTD_CCtx *c = nullptr;
    ~ZstdContexts() {
        if (d) ZSTD_freeDCtx(d);
        if (c) ZSTD_freeCCtx(c);
    }
};
thread_local ZstdContexts t_zstd;

ZSTD_DCtx *dctx() {
    if (!t_zstd.d && !(t_zstd.d = ZSTD_createDCtx())) throw std::bad_alloc();
    return t_zstd.d;
}
ZSTD_CCtx *cctx() {
    if (!t_zstd.c && !(t_zstd.c = ZSTD_createCCtx())) throw std::bad_alloc();
    return t_zstd.c;
}

// A file opened for reading (absent: fd < 0 and errno ENOENT).
struct File {
    int fd = -1;
    uint64_t size = 0;
    File() = default;
    File(const File &) = delete;
    File &operator=(const File &) = delete;
    ~File() {
        if (fd >= 0) ::close(fd);
    }
    // false if the file does not exist; KY-E110 for any other failure.
    bool open(const std::string &path) {
        fd = ::open(path.c_str(), O_RDONLY | O_CLOEXEC);
        if (fd < 0) {
            if (errno == ENOENT || errno == ENOTDIR) return false;
            io_error("open", path, std::strerror(errno));
        }
        struct stat st;
        if (::fstat(fd, &st) != 0) io_error("stat", path, std::strerror(errno));
        size = (uint64_t)st.st_size;
        return true;
    }
    void read_at(uint64_t offset, void *dst, uint64_t n, const std::string &path) const {
        uint8_t *p = (uint8_t *)dst;
        while (n) {
            ssize_t got = ::pread(fd, p, (size_t)std::min<uint64_t>(n, 1ull << 30), (off_t)offset);
            if (got < 0) {
                if (errno == EINTR) continue;
                io_error("read", path, std::strerror(errno));
            }
            if (got == 0) io_error("read", path, "the file ends early");
            p += got;
            offset += (uint64_t)got;
            n -= (uint64_t)got;
        }
    }
};

bool read_whole(const std::string &path, std::string &out) {
    File f;
    if (!f.open(path)) return false;
    out.resize(f.size);
    if (f.size) f.read_at(0, &out[0], f.size, path);
    return true;
}

bool is_dir(const std::string &path) {
    struct stat st;
    return ::stat(path.c_str(), &st) == 0 && S_ISDIR(st.st_mode);
}

void make_dirs(const std::string &path) {
    std::string cur;
    size_t i = 0;
    while (i <= path.size()) {
        size_t j = path.find('/', i);
        if (j == std::string::npos) j = path.size();
        cur = path.substr(0, j);
        if (!cur.empty() && !is_dir(cur) && ::mkdir(cur.c_str(), 0777) != 0 && errno != EEXIST)
            io_error("create directory", cur, std::strerror(errno));
        i = j + 1;
    }
}

void write_file(const std::string &path, const void *data, size_t n) {
    int fd = ::open(path.c_str(), O_WRONLY | O_CREAT | O_TRUNC | O_CLOEXEC, 0666);
    if (fd < 0) io_error("create", path, std::strerror(errno));
    const uint8_t *p = (const uint8_t *)data;
    while (n) {
        ssize_t put = ::write(fd, p, std::min<size_t>(n, 1u << 30));
        if (put < 0) {
            if (errno == EINTR) continue;
            int e = errno;
            ::close(fd);
            io_error("write", path, std::strerror(e));
        }
        p += put;
        n -= (size_t)put;
    }
    if (::close(fd) != 0) io_error("close", path, std::strerror(errno));
}

bool all_zero(const void *p, size_t n) {
    const uint8_t *b = (const uint8_t *)p;
    size_t i = 0;
    for (; i + 8 <= n; i += 8) {
        uint64_t w;
        std::memcpy(&w, b + i, 8);
        if (w) return false;
    }
    for (; i < n; i++)
        if (b[i]) return false;
    return true;
}

std::vector<uint8_t> zstd_compress(const void *src, size_t n, int level) {
    std::vector<uint8_t> out(ZSTD_compressBound(n));
    size_t r = ZSTD_compressCCtx(cctx(), out.data(), out.size(), src, n, level);
    if (ZSTD_isError(r)) throw Error(KY_E110, std::string("zstd compression failed: ") + ZSTD_getErrorName(r));
    out.resize(r);
    return out;
}

// ------------------------------------------------------------------- JSON --

struct Json {
    enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
    bool flag = false;
    std::string text;  // a string's value, or a number's literal
    std::vector<Json> items;
    std::vector<std::pair<std::string, Json>> fields;
    const Json *find(const char *key) const {
        if (kind != Object) return nullptr;
        for (const auto &f : fields)
            if (f.first == key) return &f.second;
        return nullptr;
    }
};

class JsonReader {
  public:
    JsonReader(const std::string &s, const std::string &where) : p_(s.data()), e_(s.data() + s.size()), where_(where) {}
    Json document() {
        Json v = value(0);
        space();
        if (p_ != e_) fail("characters after the JSON value");
        return v;
    }

  private:
    const char *p_, *e_;
    std::string where_;

    [[noreturn]] void fail(const std::string &what) { io_error("parse", where_, what); }
    void space() {
        while (p_ < e_ && (*p_ == ' ' || *p_ == '\n' || *p_ == '\r' || *p_ == '\t')) ++p_;
    }
    bool literal(const char *w) {
        size_t n = std::strlen(w);
        if ((size_t)(e_ - p_) >= n && std::memcmp(p_, w, n) == 0) {
            p_ += n;
            return true;
        }
        return false;
    }
    unsigned hex4() {
        if (e_ - p_ < 4) fail("a short \\u escape");
        unsigned v = 0;
        for (int i = 0; i < 4; i++) {
            char c = *p_++;
            v <<= 4;
            if (c >= '0' && c <= '9') v |= (unsigned)(c - '0');
            else if (c >= 'a' && c <= 'f') v |= (unsigned)(c - 'a' + 10);
            else if (c >= 'A' && c <= 'F') v |= (unsigned)(c - 'A' + 10);
            else fail("a bad \\u escape");
        }
        return v;
    }
    static void utf8(std::string &out, unsigned cp) {
        if (cp < 0x80) {
            out += (char)cp;
        } else if (cp < 0x800) {
            out += (char)(0xC0 | (cp >> 6));
            out += (char)(0x80 | (cp & 0x3F));
        } else if (cp < 0x10000) {
            out += (char)(0xE0 | (cp >> 12));
            out += (char)(0x80 | ((cp >> 6) & 0x3F));
            out += (char)(0x80 | (cp & 0x3F));
        } else {
            out += (char)(0xF0 | (cp >> 18));
            out += (char)(0x80 | ((cp >> 12) & 0x3F));
            out += (char)(0x80 | ((cp >> 6) & 0x3F));
            out += (char)(0x80 | (cp & 0x3F));
        }
    }
    std::string string() {
        ++p_;  // the opening quote
        std::string out;
        while (p_ < e_ && *p_ != '"') {
            char c = *p_++;
            if ((unsigned char)c < 0x20) fail("a control character in a string");
            if (c != '\\') {
                out += c;
                continue;
            }
            if (p_ >= e_) fail("an unterminated escape");
            char x = *p_++;
            switch (x) {
                case '"': out += '"'; break;
                case '\\': out += '\\'; break;
                case '/': out += '/'; break;
                case 'b': out += '\b'; break;
                case 'f': out += '\f'; break;
                case 'n': out += '\n'; break;
                case 'r': out += '\r'; break;
                case 't': out += '\t'; break;
                case 'u': {
                    unsigned cp = hex4();
                    if (cp >= 0xD800 && cp < 0xDC00) {
                        if (!(e_ - p_ >= 2 && p_[0] == '\\' && p_[1] == 'u')) fail("an unpaired surrogate");
                        p_ += 2;
                        unsigned lo = hex4();
                        if (lo < 0xDC00 || lo >= 0xE000) fail("an unpaired surrogate");
                        cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                    } else if (cp >= 0xDC00 && cp < 0xE000) {
                        fail("an unpaired surrogate");
                    }
                    utf8(out, cp);
                    break;
                }
                default: fail("a bad escape");
            }
        }
        if (p_ >= e_) fail("an unterminated string");
        ++p_;
        return out;
    }
    Json value(int depth) {
        if (depth > 64) fail("nesting deeper than 64");
        space();
        if (p_ >= e_) fail("an unexpected end");
        Json v;
        const char c = *p_;
        if (c == '{') {
            ++p_;
            v.kind = Json::Object;
            space();
            if (p_ < e_ && *p_ == '}') {
                ++p_;
                return v;
            }
            for (;;) {
                space();
                if (p_ >= e_ || *p_ != '"') fail("an object key expected");
                std::string key = string();
                space();
                if (p_ >= e_ || *p_ != ':') fail("':' expected");
                ++p_;
                v.fields.emplace_back(std::move(key), value(depth + 1));
                space();
                if (p_ < e_ && *p_ == ',') {
                    ++p_;
                    continue;
                }
                if (p_ < e_ && *p_ == '}') {
                    ++p_;
                    return v;
                }
                fail("',' or '}' expected");
            }
        }
        if (c == '[') {
            ++p_;
            v.kind = Json::Array;
            space();
            if (p_ < e_ && *p_ == ']') {
                ++p_;
                return v;
            }
            for (;;) {
                v.items.push_back(value(depth + 1));
                space();
                if (p_ < e_ && *p_ == ',') {
                    ++p_;
                    continue;
                }
                if (p_ < e_ && *p_ == ']') {
                    ++p_;
                    return v;
                }
                fail("',' or ']' expected");
            }
        }
        if (c == '"') {
            v.kind = Json::String;
            v.text = string();
            return v;
        }
        if (literal("true")) {
            v.kind = Json::Bool;
            v.flag = true;
            return v;
        }
        if (literal("false")) {
            v.kind = Json::Bool;
            return v;
        }
        if (literal("null")) return v;
        const char *s = p_;
        if (p_ < e_ && *p_ == '-') ++p_;
        const char *digits = p_;
        while (p_ < e_ && *p_ >= '0' && *p_ <= '9') ++p_;
        if (p_ == digits) fail("an unexpected character");
        if (p_ < e_ && *p_ == '.') {
            ++p_;
            while (p_ < e_ && *p_ >= '0' && *p_ <= '9') ++p_;
        }
        if (p_ < e_ && (*p_ == 'e' || *p_ == 'E')) {
            ++p_;
            if (p_ < e_ && (*p_ == '+' || *p_ == '-')) ++p_;
            while (p_ < e_ && *p_ >= '0' && *p_ <= '9') ++p_;
        }
        v.kind = Json::Number;
        v.text.assign(s, p_);
        return v;
    }
};

std::string quoted(const std::string &s) {
    std::string out = "\"";
    for (unsigned char c : s) {
        if (c == '"') out += "\\\"";
        else if (c == '\\') out += "\\\\";
        else if (c < 0x20) {
            char b[8];
            std::snprintf(b, sizeof b, "\\u%04x", c);
            out += b;
        } else out += (char)c;
    }
    return out + "\"";
}

// Where a check fails: the error code and the start of its message.
struct Blame {
    int code;
    std::string prefix;
    [[noreturn]] void fail(const std::string &problem) const { throw Error(code, prefix + problem); }
};

const Json &need(const Json &o, const char *key, const Blame &b) {
    const Json *v = o.find(key);
    if (!v) b.fail(std::string("'") + key + "' is missing");
    return *v;
}

// JSON integers are read as u64 (they may pass 2^53, format.md §8).
uint64_t as_u64(const Json &v, const char *what, const Blame &b) {
    if (v.kind != Json::Number || v.text.empty() || v.text.find_first_not_of("0123456789") != std::string::npos)
        b.fail(std::string(what) + " is not a non-negative integer");
    uint64_t x = 0;
    for (char c : v.text) {
        uint64_t d = (uint64_t)(c - '0');
        if (x > (~uint64_t(0) - d) / 10) b.fail(std::string(what) + " does not fit 64 bits");
        x = x * 10 + d;
    }
    return x;
}

const std::string &as_str(const Json &v, const char *what, const Blame &b) {
    if (v.kind != Json::String) b.fail(std::string(what) + " is not a string");
    return v.text;
}

std::vector<uint64_t> as_shape(const Json &v, const char *what, const Blame &b) {
    if (v.kind != Json::Array) b.fail(std::string(what) + " is not a list");
    std::vector<uint64_t> out;
    for (const Json &x : v.items) out.push_back(as_u64(x, what, b));
    return out;
}

std::string shape_text(const std::vector<uint64_t> &s) {
    std::string t = "[";
    for (size_t i = 0; i < s.size(); i++) t += (i ? ", " : "") + std::to_string(s[i]);
    return t + "]";
}

Json parse_file(const std::string &path, const Blame &missing) {
    std::string text;
    if (!read_whole(path, text)) missing.fail("'" + path + "' is missing");
    return JsonReader(text, path).document();
}

// ---------------------------------------------------- Zarr v3 array nodes --

// A Zarr v3 array with the regular chunk grid and the codecs kyto writes: bytes
// (little endian) then optionally zstd, or sharding_indexed whose inner chunks use
// those and whose index is bytes (little endian) then optionally crc32c.
struct ZArray {
    std::string dir;
    std::vector<uint64_t> shape, chunk;  // chunk: the stored chunk (a shard when sharded)
    std::string dtype;
    uint32_t item = 0;
    std::string key_prefix = "c";  // "default" key encoding; "" for "v2"
    char sep = '/';
    bool zstd = false;  // on the (inner) chunks
    bool sharded = false;
    std::vector<uint64_t> inner;  // inner chunk shape
    uint64_t n_inner = 1;         // inner chunks a shard
    bool index_at_end = true, index_crc = false;

    uint64_t chunk_elems() const {
        uint64_t n = 1;
        for (uint64_t c : (sharded ? inner : chunk)) n *= c;
        return n;
    }
    std::string chunk_path(const std::vector<uint64_t> &coords) const {
        std::string p = dir + "/" + key_prefix;
        for (size_t i = 0; i < coords.size(); i++) {
            if (i || !key_prefix.empty()) p += sep;
            p += std::to_string(coords[i]);
        }
        return p;
    }
};

uint32_t dtype_size(const std::string &t) {
    if (t == "uint32" || t == "float32") return 4;
    if (t == "uint64" || t == "float64") return 8;
    return 0;
}

// [bytes(little), zstd?] of a chunk (or inner chunk).
bool parse_chunk_codecs(const Json &codecs, const Blame &b) {
    if (codecs.kind != Json::Array || codecs.items.empty()) b.fail("its codecs are not a list");
    bool zstd = false;
    for (size_t i = 0; i < codecs.items.size(); i++) {
        const Json &c = codecs.items[i];
        const std::string &name = as_str(need(c, "name", b), "a codec name", b);
        const Json *conf = c.find("configuration");
        if (i == 0) {
            if (name != "bytes") b.fail("codec '" + name + "' where 'bytes' is expected");
            const Json *endian = conf ? conf->find("endian") : nullptr;
            if (endian && as_str(*endian, "endian", b) != "little") b.fail("a big-endian bytes codec");
        } else if (name == "zstd" && !zstd) {
            zstd = true;
        } else {
            b.fail("codec '" + name + "' is not one kyto writes");
        }
    }
    return zstd;
}

ZArray open_array(const std::string &dir, const Blame &b) {
    Json j = parse_file(dir + "/zarr.json", b);
    ZArray a;
    a.dir = dir;
    if (!j.find("zarr_format") || as_u64(*j.find("zarr_format"), "zarr_format", b) != 3) b.fail("it is not Zarr v3");
    if (as_str(need(j, "node_type", b), "node_type", b) != "array") b.fail("it is not an array");
    a.shape = as_shape(need(j, "shape", b), "shape", b);
    a.dtype = as_str(need(j, "data_type", b), "data_type", b);
    a.item = dtype_size(a.dtype);
    if (!a.item) b.fail("data_type '" + a.dtype + "' is not one kyto stores");
    const Json &grid = need(j, "chunk_grid", b);
    if (as_str(need(grid, "name", b), "chunk_grid.name", b) != "regular") b.fail("its chunk grid is not regular");
    a.chunk = as_shape(need(need(grid, "configuration", b), "chunk_shape", b), "chunk_shape", b);
    if (a.chunk.size() != a.shape.size()) b.fail("its chunk shape and shape differ in rank");
    for (uint64_t c : a.chunk)
        if (c == 0) b.fail("a chunk size is 0");
    if (const Json *enc = j.find("chunk_key_encoding")) {
        const std::string &name = as_str(need(*enc, "name", b), "chunk_key_encoding.name", b);
        const Json *conf = enc->find("configuration");
        const Json *sep = conf ? conf->find("separator") : nullptr;
        if (name == "default") {
            a.key_prefix = "c";
            a.sep = sep ? as_str(*sep, "separator", b)[0] : '/';
        } else if (name == "v2") {
            a.key_prefix = "";
            a.sep = sep ? as_str(*sep, "separator", b)[0] : '.';
        } else {
            b.fail("chunk key encoding '" + name + "'");
        }
    }
    if (const Json *fill = j.find("fill_value")) {
        if (fill->kind != Json::Number || std::strtod(fill->text.c_str(), nullptr) != 0.0) b.fail("its fill_value is not 0");
    }
    const Json &codecs = need(j, "codecs", b);
    if (codecs.kind != Json::Array || codecs.items.empty()) b.fail("its codecs are not a list");
    const Json &first = codecs.items[0];
    if (as_str(need(first, "name", b), "a codec name", b) == "sharding_indexed") {
        if (codecs.items.size() != 1) b.fail("codecs follow sharding_indexed");
        const Json &conf = need(first, "configuration", b);
        a.sharded = true;
        a.inner = as_shape(need(conf, "chunk_shape", b), "the inner chunk shape", b);
        if (a.inner.size() != a.chunk.size()) b.fail("its inner chunk shape and shape differ in rank");
        for (size_t d = 0; d < a.inner.size(); d++) {
            if (a.inner[d] == 0 || a.chunk[d] % a.inner[d]) b.fail("the shard shape is not a multiple of the inner chunk shape");
            a.n_inner *= a.chunk[d] / a.inner[d];
        }
        a.zstd = parse_chunk_codecs(need(conf, "codecs", b), b);
        const Json *index = conf.find("index_codecs");
        if (index) {
            if (index->kind != Json::Array || index->items.empty()) b.fail("its index codecs are not a list");
            for (size_t i = 0; i < index->items.size(); i++) {
                const std::string &name = as_str(need(index->items[i], "name", b), "an index codec name", b);
                if (i == 0 && name == "bytes") continue;
                if (i == 1 && name == "crc32c") {
                    a.index_crc = true;
                    continue;
                }
                b.fail("index codec '" + name + "' is not one kyto writes");
            }
        }
        if (const Json *loc = conf.find("index_location")) {
            const std::string &l = as_str(*loc, "index_location", b);
            if (l != "end" && l != "start") b.fail("index_location '" + l + "'");
            a.index_at_end = l == "end";
        }
    } else {
        a.zstd = parse_chunk_codecs(codecs, b);
    }
    uint64_t bytes = a.item;
    for (uint64_t c : a.chunk) {
        if (c > MAX_BUFFERED / bytes + 1) b.fail("a chunk above 2^31 bytes");
        bytes *= c;
    }
    if (bytes > MAX_BUFFERED) b.fail("a chunk above 2^31 bytes");
    return a;
}

// Decode one stored (inner) chunk into dst, which takes exactly `want` bytes.
void decode_chunk(const ZArray &a, const uint8_t *src, size_t n, void *dst, size_t want, const std::string &where,
                  const Blame &b) {
    if (a.zstd) {
        size_t r = ZSTD_decompressDCtx(dctx(), dst, want, src, n);
        if (ZSTD_isError(r)) b.fail("a chunk of " + where + " does not decode: " + ZSTD_getErrorName(r));
        if (r != want) b.fail("a chunk of " + where + " decodes to " + std::to_string(r) + " bytes, not " + std::to_string(want));
    } else {
        if (n != want) b.fail("a chunk of " + where + " holds " + std::to_string(n) + " bytes, not " + std::to_string(want));
        std::memcpy(dst, src, want);
    }
}

// A whole chunk of an unsharded array (absent: zeros, the fill value).
void read_chunk(const ZArray &a, const std::vector<uint64_t> &coords, void *dst, const Blame &b) {
    const size_t want = (size_t)(a.chunk_elems() * a.item);
    const std::string path = a.chunk_path(coords);
    File f;
    if (!f.open(path)) {
        std::memset(dst, 0, want);
        return;
    }
    std::vector<uint8_t> raw(f.size);
    if (f.size) f.read_at(0, raw.data(), f.size, path);
    decode_chunk(a, raw.data(), raw.size(), dst, want, path, b);
}

// A shard's index: (offset, nbytes) a inner chunk, MISSING twice for an absent one;
// empty when the shard itself is absent.
std::vector<uint64_t> read_shard_index(const ZArray &a, const std::vector<uint64_t> &coords, const Blame &b) {
    const std::string path = a.chunk_path(coords);
    File f;
    if (!f.open(path)) return {};
    const uint64_t index_bytes = a.n_inner * 16 + (a.index_crc ? 4 : 0);
    if (f.size < index_bytes) b.fail("shard '" + path + "' is shorter than its index");
    std::vector<uint8_t> raw(index_bytes);
    const uint64_t at = a.index_at_end ? f.size - index_bytes : 0;
    f.read_at(at, raw.data(), index_bytes, path);
    if (a.index_crc) {
        uint32_t stored;
        std::memcpy(&stored, raw.data() + a.n_inner * 16, 4);
        if (crc32c(raw.data(), a.n_inner * 16) != stored) b.fail("the index checksum of shard '" + path + "' does not match");
    }
    std::vector<uint64_t> index(a.n_inner * 2);
    std::memcpy(index.data(), raw.data(), a.n_inner * 16);
    const uint64_t lo = a.index_at_end ? 0 : index_bytes, hi = a.index_at_end ? f.size - index_bytes : f.size;
    for (uint64_t i = 0; i < a.n_inner; i++) {
        const uint64_t off = index[2 * i], len = index[2 * i + 1];
        if (off == MISSING && len == MISSING) continue;
        if (off < lo || off > hi || len > hi - off) b.fail("shard '" + path + "' indexes bytes outside its data");
    }
    return index;
}

// Inner chunk `inner` of a shard into dst (exactly chunk_elems * item bytes; zeros if
// absent).
void read_inner(const ZArray &a, const std::vector<uint64_t> &coords, const std::vector<uint64_t> &index,
                uint64_t inner, void *dst, const Blame &b) {
    const size_t want = (size_t)(a.chunk_elems() * a.item);
    if (index.empty() || (index[2 * inner] == MISSING && index[2 * inner + 1] == MISSING)) {
        std::memset(dst, 0, want);
        return;
    }
    const std::string path = a.chunk_path(coords);
    File f;
    if (!f.open(path)) b.fail("shard '" + path + "' went missing");
    std::vector<uint8_t> raw(index[2 * inner + 1]);
    if (!raw.empty()) f.read_at(index[2 * inner], raw.data(), raw.size(), path);
    decode_chunk(a, raw.data(), raw.size(), dst, want, path, b);
}

// Check a 1-D entries array (indices, data) against the layer's or graph's choice of
// inner chunk K and shard S.
void check_entries_array(const ZArray &a, uint64_t nnz, const char *what, const Blame &b) {
    if (a.shape != std::vector<uint64_t>{nnz})
        b.fail(std::string(what) + " has shape " + shape_text(a.shape) + ", not [" + std::to_string(nnz) + "]");
    if (!a.sharded) b.fail(std::string(what) + " is not sharded");
}

// Entries [s, e) of a sharded 1-D array into dst, as a list of inner-chunk tasks (run
// by the caller on the host threads). `shards` holds every shard's index.
struct EntryTask {
    const ZArray *a;
    const std::vector<std::vector<uint64_t>> *shards;
    uint64_t chunk;  // inner chunk number
    uint64_t s, e;   // the part of [chunk*K, chunk*K + K) to copy
    uint8_t *dst;    // where entry s goes
};

void plan_entries(const ZArray &a, const std::vector<std::vector<uint64_t>> &shards, uint64_t s, uint64_t e,
                  void *dst, std::vector<EntryTask> &tasks) {
    if (s >= e) return;
    const uint64_t K = a.inner[0];
    for (uint64_t j = s / K; j <= (e - 1) / K; j++) {
        const uint64_t lo = std::max(s, j * K), hi = std::min(e, j * K + K);
        tasks.push_back({&a, &shards, j, lo, hi, (uint8_t *)dst + (lo - s) * a.item});
    }
}

void run_entry_task(const EntryTask &t, std::vector<uint8_t> &scratch, const Blame &b) {
    const ZArray &a = *t.a;
    const uint64_t K = a.inner[0], per = a.n_inner;
    const uint64_t shard = t.chunk / per, inner = t.chunk % per;
    const std::vector<uint64_t> coords{shard};
    const std::vector<uint64_t> &index = (*t.shards)[shard];
    if (t.s == t.chunk * K && t.e == t.chunk * K + K) {
        read_inner(a, coords, index, inner, t.dst, b);  // straight into the destination
        return;
    }
    scratch.resize(K * a.item);
    read_inner(a, coords, index, inner, scratch.data(), b);
    std::memcpy(t.dst, scratch.data() + (t.s - t.chunk * K) * a.item, (t.e - t.s) * a.item);
}

void run_entry_tasks(const std::vector<EntryTask> &tasks, const Blame &b) {
    parallel_tasks(tasks.size(), [&](size_t i) {
        thread_local std::vector<uint8_t> scratch;
        run_entry_task(tasks[i], scratch, b);
    });
}

std::vector<std::vector<uint64_t>> read_all_shard_indexes(const ZArray &a, const Blame &b) {
    const uint64_t n = a.shape[0] ? (a.shape[0] + a.chunk[0] - 1) / a.chunk[0] : 0;
    std::vector<std::vector<uint64_t>> out(n);
    parallel_tasks(n, [&](size_t i) { out[i] = read_shard_index(a, {(uint64_t)i}, b); });
    return out;
}

// A whole unsharded 1-D array (tile_offsets, a graph's indptr).
template <typename T>
std::vector<T> read_1d(const ZArray &a, const Blame &b) {
    const uint64_t n = a.shape[0], c = a.chunk[0];
    std::vector<T> out(n);
    const uint64_t chunks = (n + c - 1) / c;
    parallel_tasks(chunks, [&](size_t i) {
        std::vector<T> buf(c);
        read_chunk(a, {(uint64_t)i}, buf.data(), b);
        const uint64_t lo = i * c, hi = std::min(n, lo + c);
        std::memcpy(out.data() + lo, buf.data(), (hi - lo) * sizeof(T));
    });
    return out;
}

// --------------------------------------------------------------- the store --

// The root's format_version (format.md §3, §11) and schema.json (§8).
Json open_store(const std::string &store) {
    Blame not_store{KY_E116, "'" + store + "' is not a kyto store: "};
    std::string text;
    if (!read_whole(store + "/zarr.json", text)) not_store.fail("its root zarr.json is missing");
    Json root = JsonReader(text, store + "/zarr.json").document();
    const Json *attrs = root.find("attributes");
    const Json *kyto = attrs ? attrs->find("kyto") : nullptr;
    const Json *version = kyto ? kyto->find("format_version") : nullptr;
    if (!version) not_store.fail("its root zarr.json has no attributes.kyto.format_version");
    const std::string found = version->kind == Json::String ? version->text : std::string("?");
    if (found != "0.1")
        throw Error(KY_E115, "the store at '" + store + "' has format_version '" + found + "', but this library reads only 0.1");
    return parse_file(store + "/schema.json", Blame{KY_E110, "could not read the schema of '" + store + "': "});
}

// Member names (format.md §2.1).
void check_name(const std::string &kind, const std::string &name) {
    auto bad = [&](const std::string &rule) { throw Error(KY_E102, "invalid " + kind + " name '" + name + "': " + rule); };
    if (name.empty() || name.size() > 64) bad("it must be 1 to 64 bytes");
    for (char c : name)
        if (!(std::isalnum((unsigned char)c) || c == '_' || c == '.' || c == '-')) bad("only A-Z a-z 0-9 _ . - are allowed");
    if (name[0] == '.' || name.rfind("__", 0) == 0) bad("it must not start with '.' or '__'");
    if (name.back() == '.') bad("it must not end with '.'");
    std::string stem = name.substr(0, name.find('.'));
    for (char &c : stem) c = (char)std::toupper((unsigned char)c);
    static const char *reserved[] = {"CON", "PRN", "AUX", "NUL"};
    for (const char *r : reserved)
        if (stem == r) bad("it is a Windows device name");
    if (stem.size() == 4 && (stem.rfind("COM", 0) == 0 || stem.rfind("LPT", 0) == 0) && stem[3] >= '1' && stem[3] <= '9')
        bad("it is a Windows device name");
}

// A member of the same name, compared case-insensitively (KY-E103).
void check_absent(const std::string &group_dir, const std::string &kind, const std::string &name) {
    DIR *d = ::opendir(group_dir.c_str());
    if (!d) return;
    std::string lower = name;
    for (char &c : lower) c = (char)std::tolower((unsigned char)c);
    std::string clash;
    while (dirent *e = ::readdir(d)) {
        std::string n = e->d_name, l = n;
        for (char &c : l) c = (char)std::tolower((unsigned char)c);
        if (l == lower) clash = n;
    }
    ::closedir(d);
    if (!clash.empty()) throw Error(KY_E103, kind + " '" + name + "' already exists in the store as '" + clash + "'");
}

const Json *member_entry(const Json &schema, const char *group, const std::string &name, const std::string &store,
                         const char *kind) {
    const Json *list = schema.find(group);
    std::string known;
    if (list && list->kind == Json::Array)
        for (const Json &m : list->items) {
            const Json *n = m.find("name");
            if (n && n->kind == Json::String) {
                if (n->text == name) return &m;
                known += (known.empty() ? "" : ", ") + n->text;
            }
        }
    throw Error(KY_E109, std::string("no ") + kind + " named '" + name + "' in the store at '" + store + "'" +
                             (known.empty() ? std::string() : " (it has: " + known + ")"));
}

std::string group_json(const std::string &attributes) {
    return "{\n  \"zarr_format\": 3,\n  \"node_type\": \"group\",\n  \"attributes\": " + attributes + "\n}\n";
}

std::string codec_list(int level) {
    return "[{\"name\": \"bytes\", \"configuration\": {\"endian\": \"little\"}}, {\"name\": \"zstd\", \"configuration\": "
           "{\"level\": " + std::to_string(level) + ", \"checksum\": false}}]";
}

std::string array_json(const std::vector<uint64_t> &shape, const char *dtype, const std::vector<uint64_t> &chunk,
                       const std::string &fill, const std::string &codecs, const std::string &attributes,
                       const std::vector<std::string> &dims) {
    std::string dn = "[";
    for (size_t i = 0; i < dims.size(); i++) dn += (i ? ", " : "") + quoted(dims[i]);
    dn += "]";
    return "{\n  \"zarr_format\": 3,\n  \"node_type\": \"array\",\n  \"shape\": " + shape_text(shape) +
           ",\n  \"data_type\": \"" + dtype + "\",\n  \"chunk_grid\": {\"name\": \"regular\", \"configuration\": "
           "{\"chunk_shape\": " + shape_text(chunk) + "}},\n  \"chunk_key_encoding\": {\"name\": \"default\", "
           "\"configuration\": {\"separator\": \"/\"}},\n  \"fill_value\": " + fill + ",\n  \"codecs\": " + codecs +
           ",\n  \"attributes\": " + attributes + ",\n  \"dimension_names\": " + dn + "\n}\n";
}

std::string sharding_codecs(const std::vector<uint64_t> &inner, int level) {
    return "[{\"name\": \"sharding_indexed\", \"configuration\": {\"chunk_shape\": " + shape_text(inner) +
           ", \"codecs\": " + codec_list(level) + ", \"index_codecs\": [{\"name\": \"bytes\", \"configuration\": "
           "{\"endian\": \"little\"}}, {\"name\": \"crc32c\"}], \"index_location\": \"end\"}}]";
}

void check_options(uint64_t K, uint64_t S, int level) {
    const uint64_t max_entries = MAX_BUFFERED / 4;
    if (K < 1 || K > max_entries)
        throw Error(KY_E104, "invalid nnz_chunk_size = " + std::to_string(K) + " (between 1 and " + std::to_string(max_entries) + " entries)");
    if (S < 1 || S > max_entries || S % K)
        throw Error(KY_E104, "invalid nnz_shard_size = " + std::to_string(S) + " (a positive multiple of nnz_chunk_size of at most " +
                                 std::to_string(max_entries) + " entries)");
    if (level < 1 || level > 22) throw Error(KY_E104, "invalid zstd_level = " + std::to_string(level) + " (between 1 and 22)");
}

// An unsharded 1-D array of chunk c whole (last chunk padded with the fill value;
// chunks of fill values not stored).
void write_1d(const std::string &dir, const void *values, uint64_t n, uint32_t item, uint64_t c, int level) {
    const uint64_t chunks = (n + c - 1) / c;
    make_dirs(dir);
    parallel_tasks(chunks, [&](size_t i) {
        std::vector<uint8_t> buf(c * item, 0);
        const uint64_t lo = i * c, hi = std::min(n, lo + c);
        std::memcpy(buf.data(), (const uint8_t *)values + lo * item, (hi - lo) * item);
        if (all_zero(buf.data(), buf.size())) return;
        std::vector<uint8_t> z = zstd_compress(buf.data(), buf.size(), level);
        make_dirs(dir + "/c");
        write_file(dir + "/c/" + std::to_string(i), z.data(), z.size());
    });
}

// A sharded 1-D array: shards of S entries, inner chunks of K, index (offset, nbytes)
// pairs + crc32c at the end; inner chunks of fill values are absent, and so is a
// shard of nothing but those.
void write_sharded(const std::string &dir, const void *values, uint64_t n, uint32_t item, uint64_t K, uint64_t S,
                   int level) {
    const uint64_t shards = (n + S - 1) / S, per = S / K;
    make_dirs(dir);
    parallel_tasks(shards, [&](size_t s) {
        std::vector<uint8_t> out, chunk(K * item);
        std::vector<uint64_t> index(per * 2, MISSING);
        for (uint64_t j = 0; j < per; j++) {
            const uint64_t lo = s * S + j * K;
            if (lo >= n) break;
            const uint64_t hi = std::min(n, lo + K);
            std::fill(chunk.begin(), chunk.end(), 0);
            std::memcpy(chunk.data(), (const uint8_t *)values + lo * item, (hi - lo) * item);
            if (all_zero(chunk.data(), chunk.size())) continue;
            std::vector<uint8_t> z = zstd_compress(chunk.data(), chunk.size(), level);
            index[2 * j] = out.size();
            index[2 * j + 1] = z.size();
            out.insert(out.end(), z.begin(), z.end());
        }
        if (out.empty()) return;
        const size_t at = out.size();
        out.resize(at + per * 16 + 4);
        std::memcpy(out.data() + at, index.data(), per * 16);
        const uint32_t crc = crc32c(out.data() + at, per * 16);
        std::memcpy(out.data() + at + per * 16, &crc, 4);
        make_dirs(dir + "/c");
        write_file(dir + "/c/" + std::to_string(s), out.data(), out.size());
    });
}

}  // namespace

// ------------------------------------------------------------- the layer --

struct StoreLayer {
    std::string store, name;
    uint64_t n_cells = 0, n_genes = 0, C = 0, G = 0, ncb = 0, ngb = 0, n_tiles = 0, nnz = 0;
    bool counts = false;  // uint32 values
    ZArray indptr, indices, data;
    std::vector<uint64_t> offsets;  // tile_offsets
    std::vector<std::vector<uint64_t>> index_shards, data_shards;
    std::vector<uint32_t> tile_ptr, tmp_idx, tmp_val;  // reused between blocks
    Blame corrupt() const { return Blame{KY_E107, "layer '" + name + "' is corrupt: "}; }
};

}  // namespace ky

struct ky_layer : ky::StoreLayer {};

namespace ky {

ky_layer *layer_open(const std::string &store, const std::string &name) {
    Json schema = open_store(store);
    const Blame sb{KY_E110, "the schema of '" + store + "' is invalid: "};
    std::unique_ptr<ky_layer> L(new ky_layer());
    L->store = store;
    L->name = name;
    L->n_cells = as_u64(need(schema, "n_cells", sb), "n_cells", sb);
    L->n_genes = as_u64(need(schema, "n_genes", sb), "n_genes", sb);
    const Json &tiling = need(schema, "tiling", sb);
    L->C = as_u64(need(tiling, "cell_block_size", sb), "cell_block_size", sb);
    L->G = as_u64(need(tiling, "gene_block_size", sb), "gene_block_size", sb);
    if (L->C == 0 || L->G == 0) sb.fail("a block size is 0");
    const Json &entry = *member_entry(schema, "layers", name, store, "layer");
    const std::string dtype = as_str(need(entry, "dtype", sb), "a layer dtype", sb);
    L->nnz = as_u64(need(entry, "nnz", sb), "a layer nnz", sb);
    const Blame b = L->corrupt();
    if (dtype != "uint32" && dtype != "float32") b.fail("schema.json gives dtype '" + dtype + "'");
    L->counts = dtype == "uint32";
    L->ncb = (L->n_cells + L->C - 1) / L->C;
    L->ngb = (L->n_genes + L->G - 1) / L->G;
    L->n_tiles = L->ncb * L->ngb;
    const std::string dir = store + "/layers/" + name;
    // indptr: uint32 [n_cell_blocks, n_gene_blocks, C + 1], one shard a cell block, one
    // inner chunk a tile (format.md §4.2, §4.4).
    L->indptr = open_array(dir + "/indptr", b);
    const std::vector<uint64_t> want_shape{L->ncb, L->ngb, L->C + 1};
    if (L->indptr.dtype != "uint32" || L->indptr.shape != want_shape || !L->indptr.sharded ||
        L->indptr.chunk != std::vector<uint64_t>{1, std::max<uint64_t>(1, L->ngb), L->C + 1} ||
        L->indptr.inner != std::vector<uint64_t>{1, 1, L->C + 1})
        b.fail("indptr is not uint32 " + shape_text(want_shape) + " in shards of one cell block and chunks of one tile");
    ZArray offsets = open_array(dir + "/tile_offsets", b);
    if (offsets.dtype != "uint64" || offsets.shape != std::vector<uint64_t>{L->n_tiles + 1} || offsets.sharded)
        b.fail("tile_offsets is not an unsharded uint64 [" + std::to_string(L->n_tiles + 1) + "]");
    L->indices = open_array(dir + "/indices", b);
    L->data = open_array(dir + "/data", b);
    check_entries_array(L->indices, L->nnz, "indices", b);
    check_entries_array(L->data, L->nnz, "data", b);
    if (L->indices.dtype != "uint32") b.fail("indices are " + L->indices.dtype + ", not uint32");
    if (L->data.dtype != dtype) b.fail("data is " + L->data.dtype + " but schema.json gives " + dtype);
    // Check tile_offsets: n_tiles + 1 values from 0, non-decreasing, ending at nnz.
    L->offsets = read_1d<uint64_t>(offsets, b);
    if (L->offsets[0] != 0) b.fail("tile_offsets[0] is not 0");
    for (uint64_t t = 0; t < L->n_tiles; t++)
        if (L->offsets[t + 1] < L->offsets[t]) b.fail("tile_offsets decreases after tile " + std::to_string(t));
    if (L->offsets[L->n_tiles] != L->nnz)
        b.fail("tile_offsets ends at " + std::to_string(L->offsets[L->n_tiles]) + " but indices and data hold " +
               std::to_string(L->nnz) + " entries");
    L->index_shards = read_all_shard_indexes(L->indices, b);
    L->data_shards = read_all_shard_indexes(L->data, b);
    return L.release();
}

void layer_shape(const ky_layer &L, uint64_t *n_cells, uint64_t *n_genes, uint64_t *nnz, uint64_t *cell_block,
                 uint64_t *gene_block, bool *counts) {
    *n_cells = L.n_cells;
    *n_genes = L.n_genes;
    *nnz = L.nnz;
    *cell_block = L.C;
    *gene_block = L.G;
    *counts = L.counts;
}

void layer_close(ky_layer *L) { delete L; }

// Rows [first, first + n) as a CSR of their own (global gene indices, values as f32;
// uint32 counts converted): the tiles of their cell blocks, gene blocks in increasing
// order (format.md §4.7, §4.8; tile_decode.rs:74-130), each tile's indptr and local
// indices checked. Returns the rows' number of values; above `capacity` nothing but
// indptr is written.
int64_t layer_read_rows(ky_layer &L, uint64_t first, uint64_t n, uint64_t *ip, uint32_t *ix, float *val,
                        uint64_t capacity) {
    KY_CHECK(first <= L.n_cells && n <= L.n_cells - first, KY_E215,
             "store: rows " + std::to_string(first) + ".." + std::to_string(first + n) + " are not rows of the layer");
    ip[0] = 0;
    if (n == 0) return 0;
    const Blame b = L.corrupt();
    const uint64_t C = L.C, ngb = L.ngb, row_len = C + 1;
    const uint64_t b0 = first / C, b1 = (first + n - 1) / C, nb = b1 - b0 + 1;
    // Each block's tile indptrs (one shard a block, one inner chunk a tile).
    L.tile_ptr.resize(nb * ngb * row_len);
    std::vector<std::vector<uint64_t>> shard_index(nb);
    parallel_tasks(nb, [&](size_t k) { shard_index[k] = read_shard_index(L.indptr, {b0 + k, 0, 0}, b); });
    parallel_tasks(nb * ngb, [&](size_t t) {
        const uint64_t k = t / ngb, g = t % ngb;
        uint32_t *tp = L.tile_ptr.data() + t * row_len;
        read_inner(L.indptr, {b0 + k, 0, 0}, shard_index[k], g, tp, b);
        // The tile's indptr: from 0, non-decreasing, flat over the padding rows, ending
        // at the tile's nnz (format.md §4.3; tile_decode.rs check_tile_indptr).
        const uint64_t blk = b0 + k, tile = blk * ngb + g;
        const uint64_t rows = std::min(C, L.n_cells - blk * C);
        const uint64_t tile_nnz = L.offsets[tile + 1] - L.offsets[tile];
        const std::string at = "tile (" + std::to_string(blk) + ", " + std::to_string(g) + ")";
        if (tp[0] != 0) b.fail(at + ": indptr does not start at 0");
        for (uint64_t r = 0; r < rows; r++)
            if (tp[r + 1] < tp[r]) b.fail(at + ": indptr decreases at row " + std::to_string(r));
        for (uint64_t r = rows; r < C; r++)
            if (tp[r + 1] != tp[rows]) b.fail(at + ": indptr is not flat over the padding rows");
        if (tp[C] != tile_nnz)
            b.fail(at + ": indptr ends at " + std::to_string(tp[C]) + " but tile_offsets gives " + std::to_string(tile_nnz) + " entries");
    });
    // Row lengths and the block's indptr.
    for (uint64_t i = 0; i < n; i++) {
        const uint64_t R = first + i, k = R / C - b0, r = R % C;
        const uint32_t *tp = L.tile_ptr.data() + k * ngb * row_len;
        uint64_t len = 0;
        for (uint64_t g = 0; g < ngb; g++, tp += row_len) len += tp[r + 1] - tp[r];
        ip[i + 1] = ip[i] + len;
    }
    const uint64_t total = ip[n];
    if (total > capacity) return (int64_t)total;
    // The entries: each block's needed run [lo, hi) of indices and data. With one gene
    // block a run is already in row order and decodes straight into the output;
    // otherwise into a staging copy that the row scatter below reorders.
    std::vector<EntryTask> tasks;
    struct Run {
        uint64_t k, r0, r1, lo, hi, out, stage;
    };
    std::vector<Run> runs;
    uint64_t staged = 0;
    for (uint64_t k = 0; k < nb; k++) {
        const uint64_t blk = b0 + k;
        const uint64_t r0 = (blk == b0 ? first % C : 0);
        const uint64_t r1 = (blk == b1 ? (first + n - 1) % C + 1 : C);
        const uint32_t *tp = L.tile_ptr.data() + k * ngb * row_len;
        const uint64_t t0 = blk * ngb;
        const uint64_t lo = L.offsets[t0] + tp[r0];
        const uint64_t hi = L.offsets[t0 + ngb - 1] + tp[(ngb - 1) * row_len + r1];
        const uint64_t out = ip[blk * C + r0 - first];
        runs.push_back({k, r0, r1, lo, hi, out, staged});
        if (ngb > 1) staged += hi - lo;
    }
    if (ngb > 1) {
        L.tmp_idx.resize(staged);
        L.tmp_val.resize(staged);
    }
    for (const Run &run : runs) {
        uint32_t *di = ngb > 1 ? L.tmp_idx.data() + run.stage : ix + run.out;
        uint32_t *dv = ngb > 1 ? L.tmp_val.data() + run.stage : (uint32_t *)(val + run.out);
        plan_entries(L.indices, L.index_shards, run.lo, run.hi, di, tasks);
        plan_entries(L.data, L.data_shards, run.lo, run.hi, dv, tasks);
    }
    run_entry_tasks(tasks, b);
    // Rows: global gene indices, values as f32, each tile row's local indices strictly
    // increasing and below its block's width (tile_decode.rs check_tile_rows).
    const uint64_t per_task = 1024;
    parallel_tasks((n + per_task - 1) / per_task, [&](size_t t) {
        for (uint64_t i = t * per_task, end = std::min(n, i + per_task); i < end; i++) {
            const uint64_t R = first + i, blk = R / C, k = blk - b0, r = R % C;
            const Run &run = runs[k];
            const uint32_t *tp = L.tile_ptr.data() + k * ngb * row_len;
            uint64_t o = ip[i];
            for (uint64_t g = 0; g < ngb; g++, tp += row_len) {
                const uint64_t width = std::min(L.G, L.n_genes - g * L.G);
                const uint64_t s = L.offsets[blk * ngb + g] + tp[r], len = tp[r + 1] - tp[r];
                const uint32_t *si, *sv;
                if (ngb > 1) {
                    si = L.tmp_idx.data() + run.stage + (s - run.lo);
                    sv = L.tmp_val.data() + run.stage + (s - run.lo);
                } else {
                    si = ix + o;
                    sv = (const uint32_t *)(val + o);
                }
                const uint32_t base = (uint32_t)(g * L.G);
                for (uint64_t e = 0; e < len; e++) {
                    const uint32_t local = si[e];
                    if (local >= width || (e && local <= si[e - 1]))
                        b.fail("tile (" + std::to_string(blk) + ", " + std::to_string(g) + ") row " + std::to_string(r) +
                               ": local gene index " + std::to_string(local) +
                               (local >= width ? " is not below the block's width " + std::to_string(width)
                                               : " does not increase"));
                    ix[o + e] = base + local;
                    const uint32_t bits = sv[e];
                    if (L.counts) val[o + e] = (float)bits;
                    else std::memcpy(&val[o + e], &bits, 4);
                }
                o += len;
            }
        }
    });
    return (int64_t)total;
}

// ------------------------------------------------------ embeddings, graphs --

void write_embedding(const std::string &store, const std::string &name, int axis, const float *values, uint64_t rows,
                     uint32_t width, uint64_t chunk_rows, int level) {
    check_name("embedding", name);
    KY_CHECK(axis == 0 || axis == 1, KY_E215, "store: an embedding's axis is cells (0) or genes (1)");
    KY_CHECK(width >= 1 && chunk_rows >= 1, KY_E215, "store: an embedding needs width and chunk rows of at least 1");
    KY_CHECK(rows == 0 || values != nullptr, KY_E215, "store: embedding values are null");
    check_options(1, 1, level);
    if (chunk_rows * (uint64_t)width * 4 > MAX_BUFFERED)
        throw Error(KY_E105, "embedding chunk is " + std::to_string(chunk_rows * (uint64_t)width * 4) +
                                 " bytes, above the format limit of 2147483648");
    const std::string group = store + "/embeddings";
    check_absent(group, "embedding", name);
    const std::string dir = group + "/" + name;
    make_dirs(dir);
    const uint64_t chunks = (rows + chunk_rows - 1) / chunk_rows;
    const uint64_t elems = chunk_rows * width;
    parallel_tasks(chunks, [&](size_t i) {
        std::vector<float> buf(elems, 0.0f);
        const uint64_t lo = i * chunk_rows, hi = std::min(rows, lo + chunk_rows);
        std::memcpy(buf.data(), values + lo * width, (hi - lo) * width * sizeof(float));
        if (all_zero(buf.data(), buf.size() * 4)) return;
        std::vector<uint8_t> z = zstd_compress(buf.data(), buf.size() * 4, level);
        make_dirs(dir + "/c/" + std::to_string(i));
        write_file(dir + "/c/" + std::to_string(i) + "/0", z.data(), z.size());
    });
    const bool cells = axis == 0;
    std::string meta = array_json({rows, width}, "float32", {chunk_rows, width}, "0.0", codec_list(level),
                                  std::string("{\"kyto\": {\"axis\": \"") + (cells ? "cells" : "genes") + "\"}}",
                                  {cells ? "cell" : "gene", "component"});
    write_file(dir + "/zarr.json", meta.data(), meta.size());
    if (!read_whole(group + "/zarr.json", meta)) {
        std::string g = group_json("{}");
        write_file(group + "/zarr.json", g.data(), g.size());
    }
}

void write_graph(const std::string &store, const std::string &name, const uint64_t *indptr, const uint32_t *indices,
                 const float *weights, uint64_t n_cells, uint64_t K, uint64_t S, int level) {
    check_name("graph", name);
    check_options(K, S, level);
    KY_CHECK(indptr != nullptr, KY_E215, "store: graph indptr is null");
    const uint64_t nnz = indptr[n_cells];
    KY_CHECK(nnz == 0 || (indices && weights), KY_E215, "store: graph indices or weights are null");
    // Graph::check (graph_io.rs:71-105).
    KY_CHECK(indptr[0] == 0, KY_E215, "store: graph indptr must run from 0 to the edges");
    for (uint64_t r = 0; r < n_cells; r++) {
        KY_CHECK(indptr[r + 1] >= indptr[r], KY_E215, "store: graph row " + std::to_string(r) + ": indptr decreases");
        for (uint64_t e = indptr[r]; e < indptr[r + 1]; e++) {
            KY_CHECK(indices[e] < n_cells, KY_E215,
                     "store: graph row " + std::to_string(r) + ": index " + std::to_string(indices[e]) + " is not below the " +
                         std::to_string(n_cells) + " cells");
            KY_CHECK(e == indptr[r] || indices[e] > indices[e - 1], KY_E215,
                     "store: graph row " + std::to_string(r) + ": indices must be strictly increasing");
        }
    }
    const std::string group = store + "/graphs";
    check_absent(group, "graph", name);
    const std::string dir = group + "/" + name;
    make_dirs(dir);
    write_1d(dir + "/indptr", indptr, n_cells + 1, 8, OFFSETS_CHUNK, level);
    std::string meta = array_json({n_cells + 1}, "uint64", {OFFSETS_CHUNK}, "0", codec_list(level), "{}", {"row"});
    write_file(dir + "/indptr/zarr.json", meta.data(), meta.size());
    make_dirs(dir + "/indices");
    write_sharded(dir + "/indices", indices, nnz, 4, K, S, level);
    meta = array_json({nnz}, "uint32", {S}, "0", sharding_codecs({K}, level), "{}", {"entry"});
    write_file(dir + "/indices/zarr.json", meta.data(), meta.size());
    make_dirs(dir + "/data");
    write_sharded(dir + "/data", weights, nnz, 4, K, S, level);
    meta = array_json({nnz}, "float32", {S}, "0.0", sharding_codecs({K}, level), "{}", {"entry"});
    write_file(dir + "/data/zarr.json", meta.data(), meta.size());
    meta = group_json("{\"kyto\": {\"kind\": \"csr\", \"axis\": \"cells\"}}");
    write_file(dir + "/zarr.json", meta.data(), meta.size());
    if (!read_whole(group + "/zarr.json", meta)) {
        std::string g = group_json("{}");
        write_file(group + "/zarr.json", g.data(), g.size());
    }
}

// A tiled layer (format.md §4; tile_encode.rs, store_writer.rs): the CSR (global
// gene indices strictly increasing in a row, no stored zeros) cut into tiles of C
// cells x G genes, each tile's rows with local gene indices, tiles back to back in
// tile-number order; indptr shards of one cell block (inner chunks of one tile, empty
// tiles absent), tile_offsets in unsharded chunks of 65,536, indices and data in
// shards of S entries (inner chunks of K).
void write_layer(const std::string &store, const std::string &name, const uint64_t *indptr, const uint32_t *indices,
                 const void *values, bool counts, uint64_t n_cells, uint64_t n_genes, uint64_t C, uint64_t G, uint64_t K,
                 uint64_t S, int level) {
    check_name("layer", name);
    check_options(K, S, level);
    if (C < 1 || G < 1) throw Error(KY_E104, "invalid tile size C = " + std::to_string(C) + ", G = " + std::to_string(G));
    if (C * G > 0xFFFFFFFFull || C > 0xFFFFFFFFull || G > 0xFFFFFFFFull)
        throw Error(KY_E105, "C * G is " + std::to_string(C * G) + ", above the format limit of 4294967295");
    if (n_genes > 0xFFFFFFFFull || n_cells > 0xFFFFFFFFull)
        throw Error(KY_E105, "an axis length is above the format limit of 4294967295");
    const uint64_t ncb = (n_cells + C - 1) / C, ngb = (n_genes + G - 1) / G, row_len = C + 1;
    if (std::max<uint64_t>(1, ngb) * row_len * 4 > MAX_BUFFERED)
        throw Error(KY_E105, "one indptr shard is " + std::to_string(std::max<uint64_t>(1, ngb) * row_len * 4) +
                                 " bytes, above the format limit of 2147483648");
    KY_CHECK(indptr != nullptr, KY_E215, "store: layer indptr is null");
    const uint64_t nnz = indptr[n_cells];
    KY_CHECK(nnz == 0 || (indices && values), KY_E215, "store: layer indices or values are null");
    const std::string what = "invalid sparse matrix for layer '" + name + "': ";
    if (indptr[0] != 0) throw Error(KY_E106, what + "indptr does not start at 0");
    for (uint64_t r = 0; r < n_cells; r++)
        if (indptr[r + 1] < indptr[r]) throw Error(KY_E106, what + "indptr decreases at row " + std::to_string(r));
    // Tile sizes, a cell block a task; entries and explicit zeros checked on the way.
    std::vector<uint64_t> tile_nnz(ncb * ngb + 1, 0);
    parallel_tasks(ncb, [&](size_t b) {
        for (uint64_t r = b * C; r < std::min(n_cells, b * C + C); r++)
            for (uint64_t e = indptr[r]; e < indptr[r + 1]; e++) {
                const uint32_t j = indices[e];
                if (j >= n_genes || (e > indptr[r] && j <= indices[e - 1]))
                    throw Error(KY_E106, what + "row " + std::to_string(r) + ": gene index " + std::to_string(j) +
                                             (j >= n_genes ? " is not below n_genes" : " does not increase"));
                uint32_t bits;
                std::memcpy(&bits, (const uint8_t *)values + e * 4, 4);
                if (counts ? bits == 0 : (bits & 0x7FFFFFFFu) == 0)
                    throw Error(KY_E106, what + "row " + std::to_string(r) + " stores a zero (format.md 4.6 forbids it)");
                tile_nnz[b * ngb + j / G]++;
            }
    });
    std::vector<uint64_t> offsets(ncb * ngb + 1, 0);
    for (uint64_t t = 0; t < ncb * ngb; t++) offsets[t + 1] = offsets[t] + tile_nnz[t];
    // Entries in tile order with local indices; the tiles' indptrs a block at a time.
    std::vector<uint32_t> local(nnz), vals(nnz);
    std::vector<uint32_t> tile_ptr(ncb * ngb * row_len, 0);
    parallel_tasks(ncb, [&](size_t b) {
        std::vector<uint64_t> cursor(ngb);
        for (uint64_t g = 0; g < ngb; g++) cursor[g] = offsets[b * ngb + g];
        const uint64_t rows = std::min(C, n_cells - b * C);
        for (uint64_t r = 0; r < rows; r++) {
            const uint64_t R = b * C + r;
            for (uint64_t e = indptr[R]; e < indptr[R + 1]; e++) {
                const uint64_t g = indices[e] / G;
                local[cursor[g]] = (uint32_t)(indices[e] - g * G);
                std::memcpy(&vals[cursor[g]], (const uint8_t *)values + e * 4, 4);
                cursor[g]++;
            }
            for (uint64_t g = 0; g < ngb; g++)
                tile_ptr[(b * ngb + g) * row_len + r + 1] = (uint32_t)(cursor[g] - offsets[b * ngb + g]);
        }
        for (uint64_t g = 0; g < ngb; g++) {
            uint32_t *tp = tile_ptr.data() + (b * ngb + g) * row_len;
            for (uint64_t r = rows; r < C; r++) tp[r + 1] = tp[rows];  // flat over the padding rows
        }
    });
    const std::string group = store + "/layers";
    check_absent(group, "layer", name);
    const std::string dir = group + "/" + name;
    // indptr: one shard a cell block, inner chunks of one tile (empty tiles absent).
    const uint64_t per = std::max<uint64_t>(1, ngb);
    make_dirs(dir + "/indptr");
    parallel_tasks(ngb ? ncb : 0, [&](size_t b) {
        std::vector<uint8_t> out;
        std::vector<uint64_t> index(per * 2, MISSING);
        for (uint64_t g = 0; g < ngb; g++) {
            const uint32_t *tp = tile_ptr.data() + (b * ngb + g) * row_len;
            if (all_zero(tp, row_len * 4)) continue;
            std::vector<uint8_t> z = zstd_compress(tp, row_len * 4, level);
            index[2 * g] = out.size();
            index[2 * g + 1] = z.size();
            out.insert(out.end(), z.begin(), z.end());
        }
        if (out.empty()) return;
        const size_t at = out.size();
        out.resize(at + per * 16 + 4);
        std::memcpy(out.data() + at, index.data(), per * 16);
        const uint32_t crc = crc32c(out.data() + at, per * 16);
        std::memcpy(out.data() + at + per * 16, &crc, 4);
        make_dirs(dir + "/indptr/c/" + std::to_string(b) + "/0");
        write_file(dir + "/indptr/c/" + std::to_string(b) + "/0/0", out.data(), out.size());
    });
    std::string meta = array_json({ncb, ngb, row_len}, "uint32", {1, per, row_len}, "0",
                                  sharding_codecs({1, 1, row_len}, level), "{}", {"cell_block", "gene_block", "row"});
    write_file(dir + "/indptr/zarr.json", meta.data(), meta.size());
    make_dirs(dir + "/tile_offsets");
    write_1d(dir + "/tile_offsets", offsets.data(), offsets.size(), 8, OFFSETS_CHUNK, level);
    meta = array_json({offsets.size()}, "uint64", {OFFSETS_CHUNK}, "0", codec_list(level), "{}", {"tile"});
    write_file(dir + "/tile_offsets/zarr.json", meta.data(), meta.size());
    make_dirs(dir + "/indices");
    write_sharded(dir + "/indices", local.data(), nnz, 4, K, S, level);
    meta = array_json({nnz}, "uint32", {S}, "0", sharding_codecs({K}, level), "{}", {"entry"});
    write_file(dir + "/indices/zarr.json", meta.data(), meta.size());
    make_dirs(dir + "/data");
    write_sharded(dir + "/data", vals.data(), nnz, 4, K, S, level);
    meta = array_json({nnz}, counts ? "uint32" : "float32", {S}, counts ? "0" : "0.0", sharding_codecs({K}, level), "{}",
                      {"entry"});
    write_file(dir + "/data/zarr.json", meta.data(), meta.size());
    meta = group_json("{\"kyto\": {\"kind\": \"tiled_csr\"}}");
    write_file(dir + "/zarr.json", meta.data(), meta.size());
    if (!read_whole(group + "/zarr.json", meta)) {
        std::string g = group_json("{}");
        write_file(group + "/zarr.json", g.data(), g.size());
    }
}

// An embedding of the store, checked against schema.json (embedding_io.rs
// read_embedding): shape [rows, width], float32, axis attribute.
void read_embedding(const std::string &store, const std::string &name, uint64_t *rows, uint32_t *width, float *out,
                    uint64_t capacity) {
    Json schema = open_store(store);
    const Json &entry = *member_entry(schema, "embeddings", name, store, "embedding");
    const Blame b{KY_E126, "embedding '" + name + "' is corrupt: "};
    const std::string axis = as_str(need(entry, "axis", b), "axis", b);
    const uint64_t w = as_u64(need(entry, "width", b), "width", b);
    const std::string dtype = as_str(need(entry, "dtype", b), "dtype", b);
    const Blame sb{KY_E110, "the schema of '" + store + "' is invalid: "};
    const uint64_t n = as_u64(need(schema, axis == "genes" ? "n_genes" : "n_cells", sb), "an axis length", sb);
    const std::string dir = store + "/embeddings/" + name;
    ZArray a = open_array(dir, b);
    Json meta = parse_file(dir + "/zarr.json", b);
    const Json *attrs = meta.find("attributes");
    const Json *kyto = attrs ? attrs->find("kyto") : nullptr;
    const Json *ax = kyto ? kyto->find("axis") : nullptr;
    if (!ax || ax->kind != Json::String || ax->text != axis) b.fail("attributes.kyto.axis disagrees with schema.json");
    if (a.shape != std::vector<uint64_t>{n, w}) b.fail("the array has shape " + shape_text(a.shape) + ", but schema.json gives [" +
                                                         std::to_string(n) + ", " + std::to_string(w) + "]");
    if (a.dtype != dtype || dtype != "float32") b.fail("the array has dtype " + a.dtype + "; this reader takes float32");
    if (a.sharded || a.chunk[1] != w) b.fail("its chunks are not whole rows");
    KY_CHECK(w < (1ull << 32), KY_E215, "store: an embedding wider than 2^32");
    *rows = n;
    *width = (uint32_t)w;
    if (!out) return;
    KY_CHECK(capacity >= n * w, KY_E215, "store: the output holds fewer than rows x width values");
    const uint64_t cr = a.chunk[0], chunks = (n + cr - 1) / cr;
    parallel_tasks(chunks, [&](size_t i) {
        std::vector<float> buf(cr * w);
        read_chunk(a, {(uint64_t)i, 0}, buf.data(), b);
        const uint64_t lo = i * cr, hi = std::min(n, lo + cr);
        std::memcpy(out + lo * w, buf.data(), (hi - lo) * w * sizeof(float));
    });
}

struct GraphFile {
    uint64_t n_cells = 0, nnz = 0;
    ZArray indptr, indices, data;
};

GraphFile open_graph(const std::string &store, const std::string &name) {
    Json schema = open_store(store);
    const Json &entry = *member_entry(schema, "graphs", name, store, "graph");
    const Blame b{KY_E126, "graph '" + name + "' is corrupt: "};
    const Blame sb{KY_E110, "the schema of '" + store + "' is invalid: "};
    GraphFile g;
    g.n_cells = as_u64(need(schema, "n_cells", sb), "n_cells", sb);
    g.nnz = as_u64(need(entry, "nnz", b), "nnz", b);
    const std::string dtype = as_str(need(entry, "dtype", b), "dtype", b);
    const std::string dir = store + "/graphs/" + name;
    Json meta = parse_file(dir + "/zarr.json", b);
    const Json *attrs = meta.find("attributes");
    const Json *kyto = attrs ? attrs->find("kyto") : nullptr;
    const Json *kind = kyto ? kyto->find("kind") : nullptr;
    if (!kind || kind->kind != Json::String || kind->text != "csr") b.fail("attributes.kyto.kind is not \"csr\"");
    g.indptr = open_array(dir + "/indptr", b);
    g.indices = open_array(dir + "/indices", b);
    g.data = open_array(dir + "/data", b);
    if (g.indptr.dtype != "uint64" || g.indptr.shape != std::vector<uint64_t>{g.n_cells + 1} || g.indptr.sharded)
        b.fail("indptr is not an unsharded uint64 [n_cells + 1]");
    check_entries_array(g.indices, g.nnz, "indices", b);
    check_entries_array(g.data, g.nnz, "data", b);
    if (g.indices.dtype != "uint32") b.fail("indices are not uint32");
    if (g.data.dtype != dtype || dtype != "float32") b.fail("data is " + g.data.dtype + "; this reader takes float32");
    return g;
}

void read_graph(const std::string &store, const std::string &name, uint64_t *n_cells, uint64_t *nnz, uint64_t *indptr,
                uint32_t *indices, float *weights) {
    GraphFile g = open_graph(store, name);
    *n_cells = g.n_cells;
    *nnz = g.nnz;
    if (!indptr) return;
    const Blame b{KY_E126, "graph '" + name + "' is corrupt: "};
    std::vector<uint64_t> ip = read_1d<uint64_t>(g.indptr, b);
    if (ip[0] != 0 || ip[g.n_cells] != g.nnz) b.fail("indptr does not run from 0 to the " + std::to_string(g.nnz) + " edges");
    for (uint64_t r = 0; r < g.n_cells; r++)
        if (ip[r + 1] < ip[r]) b.fail("row " + std::to_string(r) + ": indptr decreases");
    std::memcpy(indptr, ip.data(), ip.size() * 8);
    if (!indices || !weights) return;
    std::vector<EntryTask> tasks;
    auto idx_shards = read_all_shard_indexes(g.indices, b);
    auto dat_shards = read_all_shard_indexes(g.data, b);
    plan_entries(g.indices, idx_shards, 0, g.nnz, indices, tasks);
    plan_entries(g.data, dat_shards, 0, g.nnz, weights, tasks);
    run_entry_tasks(tasks, b);
    for (uint64_t r = 0; r < g.n_cells; r++)
        for (uint64_t e = ip[r]; e < ip[r + 1]; e++) {
            if (indices[e] >= g.n_cells) b.fail("row " + std::to_string(r) + ": an index is not below n_cells");
            if (e > ip[r] && indices[e] <= indices[e - 1]) b.fail("row " + std::to_string(r) + ": indices do not increase");
        }
}

}  // namespace ky
