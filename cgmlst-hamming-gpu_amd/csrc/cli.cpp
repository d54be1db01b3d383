// cli.cpp -- cgmlst-hamming-gpu: drop-in command line for the MI355X library.
//
// Same flags, defaults, output bytes, exit codes and error texts as the
// reference host (tensorcore/rs/src/main.rs:19-82 flags, 267-491 main):
// allele-call TSV in, full / lower-tri / upper-tri u16 distance TSV out (gzip
// with -z or *.gz).  Parsing and the completeness filters run on the host;
// the encoder runs on the device (cgh_dataset_encode; the host encoder and
// cgh_dataset_build when n x L reaches 2^31 calls), and the distances stream
// back stripe by stripe through cgh_dataset_stream, which keeps the callback
// contract of the reference's tch_run (main.rs:84-93) while reading the rare
// pairs from the dataset's groups instead of a per-pair CSR.

#include <errno.h>
#include <omp.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>
#include <zlib.h>

#include <chrono>
#include <string>
#include <memory>
#include <vector>

#include "cgmlst_hamming.h"

namespace {

enum class Layout { Full, LowerTri, UpperTri };

struct Args {
    std::string input;
    bool have_input = false;
    std::string output;
    bool have_output = false;
    size_t threads = 0;
    Layout layout = Layout::Full;
    std::string input_sep = "\t";
    std::string output_sep = "\t";
    std::string index_name = "cgmlst-hamming-gpu";
    std::string missing = "-";
    bool skip_repl = false;
    bool have_lc = false, have_sc = false;
    double lc = 0, sc = 0;
    bool gzip = false;
    int32_t max_dist = 0;
    size_t tau = 0;
    size_t stripe = 512;
    size_t device = 0;
    bool silent = false;
};

const char* VERSION = "0.1.0";

[[noreturn]] void die(const std::string& m) {
    fprintf(stderr, "error: %s\n", m.c_str());
    exit(1);
}

[[noreturn]] void usage_error(const std::string& m) {
    fprintf(stderr, "error: %s\n\nUsage: cgmlst-hamming-gpu [OPTIONS] --input <INPUT>\n\nFor more information, try '--help'.\n",
            m.c_str());
    exit(2);
}

void print_help() {
    printf(
        "Exact cgMLST Hamming distance matrix on MI355X matrix cores\n\n"
        "Usage: cgmlst-hamming-gpu [OPTIONS] --input <INPUT>\n\n"
        "Options:\n"
        "  -i, --input <INPUT>                Input allele-call TSV (chewBBACA format). Use '-' for stdin\n"
        "  -o, --output <OUTPUT>              Output file [default: stdout]\n"
        "  -t, --threads <THREADS>            Number of CPU threads [default: all cores] [default: 0]\n"
        "  -m, --matrix-format <MATRIX_FORMAT>  Output matrix layout [default: full] [possible values: full, lower-tri, upper-tri]\n"
        "  -d, --input-sep <INPUT_SEP>        Input field separator [default: \"\\t\"]\n"
        "  -D, --output-sep <OUTPUT_SEP>      Output field separator [default: \"\\t\"]\n"
        "  -x, --index-name <INDEX_NAME>      Name for the top-left index cell [default: cgmlst-hamming-gpu]\n"
        "  -n, --missing-char <MISSING_CHAR>  Character used for missing data [default: -]\n"
        "  -r, --skip-input-replacements      Do not strip the chewBBACA `INF-` prefix\n"
        "  -L, --locus-completeness <LOCUS_COMPLETENESS>    Keep only loci with at least this %% of non-missing calls (0-100)\n"
        "  -S, --sample-completeness <SAMPLE_COMPLETENESS>  Keep only samples with at least this %% of non-missing calls (0-100)\n"
        "  -z, --gzip                         gzip-compress the output (also enabled when -o ends with .gz)\n"
        "  -X, --max-dist <MAX_DIST>          Cap distances at this value (0 = no cap) [default: 0]\n"
        "      --tau <TAU>                    Frequency threshold: alleles seen >= tau times go to the MFMA GEMM [0 = auto] [default: 0]\n"
        "      --stripe <STRIPE>              Rows per GPU stripe [default: 512]\n"
        "      --device <DEVICE>              GPU device [default: 0]\n"
        "  -s, --silent                       Suppress progress/log messages on stderr\n"
        "  -h, --help                         Print help\n"
        "  -V, --version                      Print version\n");
}

bool parse_usize(const std::string& s, size_t* v) {
    if (s.empty()) return false;
    size_t x = 0;
    for (char c : s) {
        if (c < '0' || c > '9') return false;
        if (x > (SIZE_MAX - 9) / 10) return false;
        x = x * 10 + (size_t)(c - '0');
    }
    *v = x;
    return true;
}

bool parse_i32(const std::string& s, int32_t* v) {
    if (s.empty()) return false;
    size_t i = 0;
    bool neg = false;
    if (s[0] == '-' || s[0] == '+') {
        neg = s[0] == '-';
        i = 1;
    }
    if (i >= s.size()) return false;
    int64_t x = 0;
    for (; i < s.size(); i++) {
        if (s[i] < '0' || s[i] > '9') return false;
        x = x * 10 + (s[i] - '0');
        if (x > 2147483648LL) return false;
    }
    if (neg) x = -x;
    if (x > 2147483647LL || x < -2147483648LL) return false;
    *v = (int32_t)x;
    return true;
}

bool parse_f64(const std::string& s, double* v) {
    if (s.empty()) return false;
    char* end = NULL;
    errno = 0;
    const double x = strtod(s.c_str(), &end);
    if (*end || end == s.c_str()) return false;
    *v = x;
    return true;
}

Args parse_args(int argc, char** argv) {
    Args a;
    struct Opt {
        char s;
        const char* l;
        bool takes;
        const char* meta;
    };
    static const Opt opts[] = {
        {'i', "input", true, "--input <INPUT>"},
        {'o', "output", true, "--output <OUTPUT>"},
        {'t', "threads", true, "--threads <THREADS>"},
        {'m', "matrix-format", true, "--matrix-format <MATRIX_FORMAT>"},
        {'d', "input-sep", true, "--input-sep <INPUT_SEP>"},
        {'D', "output-sep", true, "--output-sep <OUTPUT_SEP>"},
        {'x', "index-name", true, "--index-name <INDEX_NAME>"},
        {'n', "missing-char", true, "--missing-char <MISSING_CHAR>"},
        {'r', "skip-input-replacements", false, "--skip-input-replacements"},
        {'L', "locus-completeness", true, "--locus-completeness <LOCUS_COMPLETENESS>"},
        {'S', "sample-completeness", true, "--sample-completeness <SAMPLE_COMPLETENESS>"},
        {'z', "gzip", false, "--gzip"},
        {'X', "max-dist", true, "--max-dist <MAX_DIST>"},
        {0, "tau", true, "--tau <TAU>"},
        {0, "stripe", true, "--stripe <STRIPE>"},
        {0, "device", true, "--device <DEVICE>"},
        {'s', "silent", false, "--silent"},
        {'h', "help", false, "--help"},
        {'V', "version", false, "--version"},
    };
    auto apply = [&](const Opt& o, const std::string& v) {
        const std::string name = o.l;
        auto bad = [&]() { usage_error("invalid value '" + v + "' for '" + o.meta + "'"); };
        if (name == "input") { a.input = v; a.have_input = true; }
        else if (name == "output") { a.output = v; a.have_output = true; }
        else if (name == "threads") { if (!parse_usize(v, &a.threads)) bad(); }
        else if (name == "matrix-format") {
            if (v == "full") a.layout = Layout::Full;
            else if (v == "lower-tri") a.layout = Layout::LowerTri;
            else if (v == "upper-tri") a.layout = Layout::UpperTri;
            else usage_error("invalid value '" + v + "' for '" + o.meta + "'\n  [possible values: full, lower-tri, upper-tri]");
        }
        else if (name == "input-sep") a.input_sep = v;
        else if (name == "output-sep") a.output_sep = v;
        else if (name == "index-name") a.index_name = v;
        else if (name == "missing-char") a.missing = v;
        else if (name == "skip-input-replacements") a.skip_repl = true;
        else if (name == "locus-completeness") { if (!parse_f64(v, &a.lc)) bad(); a.have_lc = true; }
        else if (name == "sample-completeness") { if (!parse_f64(v, &a.sc)) bad(); a.have_sc = true; }
        else if (name == "gzip") a.gzip = true;
        else if (name == "max-dist") { if (!parse_i32(v, &a.max_dist)) bad(); }
        else if (name == "tau") { if (!parse_usize(v, &a.tau)) bad(); }
        else if (name == "stripe") { if (!parse_usize(v, &a.stripe)) bad(); }
        else if (name == "device") { if (!parse_usize(v, &a.device)) bad(); }
        else if (name == "silent") a.silent = true;
        else if (name == "help") { print_help(); exit(0); }
        else if (name == "version") { printf("cgmlst-hamming-gpu %s\n", VERSION); exit(0); }
    };
    for (int k = 1; k < argc; k++) {
        std::string t = argv[k];
        if (t.size() > 2 && t[0] == '-' && t[1] == '-') {
            std::string name = t.substr(2), val;
            bool has_val = false;
            const size_t eq = name.find('=');
            if (eq != std::string::npos) {
                val = name.substr(eq + 1);
                name = name.substr(0, eq);
                has_val = true;
            }
            const Opt* o = NULL;
            for (auto& x : opts)
                if (name == x.l) o = &x;
            if (!o) usage_error("unexpected argument '" + t + "' found");
            if (o->takes) {
                if (!has_val) {
                    if (k + 1 >= argc) usage_error(std::string("a value is required for '") + o->meta + "' but none was supplied");
                    val = argv[++k];
                }
                apply(*o, val);
            } else {
                if (has_val) usage_error("unexpected value '" + val + "' for '" + o->meta + "' found; no more were expected");
                apply(*o, "");
            }
        } else if (t.size() > 1 && t[0] == '-' && t != "--") {
            for (size_t p = 1; p < t.size(); p++) {
                const Opt* o = NULL;
                for (auto& x : opts)
                    if (x.s && x.s == t[p]) o = &x;
                if (!o) usage_error(std::string("unexpected argument '-") + t[p] + "' found");
                if (!o->takes) {
                    apply(*o, "");
                    continue;
                }
                std::string val = t.substr(p + 1);
                if (!val.empty() && val[0] == '=') val = val.substr(1);
                if (val.empty() && p + 1 == t.size()) {
                    if (k + 1 >= argc) usage_error(std::string("a value is required for '") + o->meta + "' but none was supplied");
                    val = argv[++k];
                }
                apply(*o, val);
                break;
            }
        } else {
            usage_error("unexpected argument '" + t + "' found");
        }
    }
    if (!a.have_input) usage_error("the following required arguments were not provided:\n  --input <INPUT>");
    return a;
}

double now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Rust str::trim(): strip Unicode White_Space from both ends (UTF-8 input).
bool ws_at(const char* s, size_t len, size_t pos, size_t* clen) {
    const unsigned char c = (unsigned char)s[pos];
    if (c == ' ' || (c >= 0x09 && c <= 0x0d)) { *clen = 1; return true; }
    if (c == 0xc2 && pos + 1 < len) {
        const unsigned char d = (unsigned char)s[pos + 1];
        if (d == 0x85 || d == 0xa0) { *clen = 2; return true; }
    }
    if ((c == 0xe1 || c == 0xe2 || c == 0xe3) && pos + 2 < len) {
        const unsigned char d = (unsigned char)s[pos + 1], e = (unsigned char)s[pos + 2];
        const uint32_t cp = ((uint32_t)(c & 0x0f) << 12) | ((uint32_t)(d & 0x3f) << 6) | (e & 0x3f);
        if (cp == 0x1680 || (cp >= 0x2000 && cp <= 0x200a) || cp == 0x2028 || cp == 0x2029 || cp == 0x202f ||
            cp == 0x205f || cp == 0x3000) {
            *clen = 3;
            return true;
        }
    }
    return false;
}

size_t utf8_start_of_last(const char* s, size_t end) {
    size_t p = end - 1;
    while (p > 0 && ((unsigned char)s[p] & 0xc0) == 0x80) p--;
    return p;
}

void trim(const char*& b, const char*& e) {
    size_t cl;
    while (b < e && ws_at(b, (size_t)(e - b), 0, &cl)) b += cl;
    while (b < e) {
        const size_t p = utf8_start_of_last(b, (size_t)(e - b));
        if (ws_at(b, (size_t)(e - b), p, &cl) && p + cl == (size_t)(e - b)) e = b + p;
        else break;
    }
}

// main.rs:97-113.  Returns false on u64 overflow.
bool parse_allele(const char* b, const char* e, bool skip_repl, const std::string& missing, uint64_t* out) {
    trim(b, e);
    if (!skip_repl && e - b >= 4 && memcmp(b, "INF-", 4) == 0) b += 4;
    const size_t len = (size_t)(e - b);
    if ((len == missing.size() && memcmp(b, missing.data(), len) == 0) || len == 0) {
        *out = 0;
        return true;
    }
    for (const char* p = b; p < e; p++)
        if (*p < '0' || *p > '9') {
            *out = 0;
            return true;
        }
    uint64_t x = 0;
    for (const char* p = b; p < e; p++) {
        const uint64_t d = (uint64_t)(*p - '0');
        if (x > (UINT64_MAX - d) / 10) return false;
        x = x * 10 + d;
    }
    *out = x;
    return true;
}

bool valid_utf8(const std::string& s) {
    const unsigned char* p = (const unsigned char*)s.data();
    const size_t n = s.size();
    size_t i = 0;
    while (i < n) {
        const unsigned char c = p[i];
        if (c < 0x80) { i++; continue; }
        size_t len;
        uint32_t cp;
        if ((c & 0xe0) == 0xc0) { len = 2; cp = c & 0x1f; }
        else if ((c & 0xf0) == 0xe0) { len = 3; cp = c & 0x0f; }
        else if ((c & 0xf8) == 0xf0) { len = 4; cp = c & 0x07; }
        else return false;
        if (i + len > n) return false;
        for (size_t k = 1; k < len; k++) {
            if ((p[i + k] & 0xc0) != 0x80) return false;
            cp = (cp << 6) | (p[i + k] & 0x3f);
        }
        if ((len == 2 && cp < 0x80) || (len == 3 && cp < 0x800) || (len == 4 && (cp < 0x10000 || cp > 0x10ffff)) ||
            (cp >= 0xd800 && cp <= 0xdfff))
            return false;
        i += len;
    }
    return true;
}

// MemAvailable of /proc/meminfo in bytes, 0 when unknown.
double host_available_bytes() {
    FILE* f = fopen("/proc/meminfo", "r");
    if (!f) return 0;
    char line[256];
    double kb = 0;
    while (fgets(line, sizeof line, f))
        if (sscanf(line, "MemAvailable: %lf kB", &kb) == 1) break;
    fclose(f);
    return kb * 1024.0;
}

std::string io_error(int e) { return std::string(strerror(e)) + " (os error " + std::to_string(e) + ")"; }

bool read_all(const std::string& path, std::string* out, int* eno) {
    FILE* f = path == "-" ? stdin : fopen(path.c_str(), "rb");
    if (!f) {
        *eno = errno;
        return false;
    }
    char buf[1 << 16];
    size_t r;
    while ((r = fread(buf, 1, sizeof buf, f)) > 0) out->append(buf, r);
    const bool ok = !ferror(f);
    *eno = errno;
    if (f != stdin) fclose(f);
    return ok;
}

// The symmetric matrix as it arrives: row i of stripe s is copied in (its
// strict upper part) and mirrored into column i of the later rows, so that
// every row is contiguous by the time it is formatted.
struct Matrix {
    size_t n = 0;
    std::unique_ptr<uint16_t[]> m;  // n x n; pages are touched first by the (parallel) stripe copies
    void init(size_t rows) {
        n = rows;
        m.reset(new uint16_t[std::max<size_t>(n * n, 1)]);
    }
    void zero() {
#pragma omp parallel for schedule(static)
        for (long i = 0; i < (long)n; i++) memset(row((size_t)i), 0, n * 2);
    }
    uint16_t* row(size_t i) { return m.get() + i * n; }
    const uint16_t* row(size_t i) const { return m.get() + i * n; }
};

// Distance -> "<sep><decimal>" in 16-byte slots (sep of at most 10 bytes),
// so a cell is one table copy.
struct CellText {
    std::vector<char> slot;   // 16 bytes per value
    std::vector<uint8_t> len;
    std::string osep;
    bool table = false;
    void build(const std::string& sep, uint32_t maxv) {
        osep = sep;
        table = sep.size() <= 10;
        if (!table) return;
        slot.assign((size_t)(maxv + 1) * 16, 0);
        len.assign((size_t)maxv + 1, 0);
        for (uint32_t v = 0; v <= maxv; v++) {
            char d[12];
            const int k = snprintf(d, sizeof d, "%u", v);
            memcpy(&slot[(size_t)v * 16], sep.data(), sep.size());
            memcpy(&slot[(size_t)v * 16 + sep.size()], d, (size_t)k);
            len[v] = (uint8_t)(sep.size() + (size_t)k);
        }
    }
};

char* fmt_u(char* p, uint32_t v) {
    char tmp[12];
    int k = 0;
    do {
        tmp[k++] = (char)('0' + v % 10);
        v /= 10;
    } while (v);
    while (k) *p++ = tmp[--k];
    return p;
}

// The output text, written row by row in order (main.rs:225-265).  Row i
// needs H[i][j] for every j: the upper part of rows j <= i, so a row can be
// written as soon as its own stripe has arrived.  Rows are formatted in
// parallel in groups of GROUP; with gzip each group becomes one gzip member
// (level 3), compressed in parallel too: concatenated members are one valid
// gzip stream.  The file is opened at the first write.
struct Writer {
    static constexpr size_t GROUP = 16;
    const std::vector<std::string>* ids = NULL;
    const Matrix* mx = NULL;
    CellText cells;
    std::string osep, index_name, path;
    Layout layout = Layout::Full;
    int32_t cap = 0;
    bool gz = false, to_file = false;
    FILE* f = NULL;
    size_t next = 0;  // rows written

    void put(const char* p, size_t len) {
        if (len && fwrite(p, 1, len, f) != len) die("writing output: " + io_error(errno));
    }

    void open() {
        if (f) return;
        if (to_file) {
            f = fopen(path.c_str(), "wb");
            if (!f) die("creating " + path + ": " + io_error(errno));
        } else {
            f = stdout;
        }
        setvbuf(f, NULL, _IOFBF, 16 << 20);
        std::string head = index_name;
        for (auto& id : *ids) {
            head += osep;
            head += id;
        }
        head += "\n";
        emit({head});
    }

    // Text pieces in order -> the file (one gzip member for all of them).
    void emit(const std::vector<std::string>& parts) {
        if (!gz) {
            for (auto& t : parts) put(t.data(), t.size());
            return;
        }
        std::string z = deflate_member(parts);
        put(z.data(), z.size());
    }

    static std::string deflate_member(const std::vector<std::string>& parts) {
        z_stream zs;
        memset(&zs, 0, sizeof zs);
        if (deflateInit2(&zs, 3, Z_DEFLATED, 15 + 16, 8, Z_DEFAULT_STRATEGY) != Z_OK) die("gzip: deflateInit2");
        size_t in = 0;
        for (auto& t : parts) in += t.size();
        std::string out(deflateBound(&zs, (uLong)in) + parts.size() * 16 + 64, '\0');
        zs.next_out = (Bytef*)&out[0];
        zs.avail_out = (uInt)out.size();
        for (size_t k = 0; k < parts.size(); k++) {
            zs.next_in = (Bytef*)parts[k].data();
            zs.avail_in = (uInt)parts[k].size();
            const int flush = k + 1 == parts.size() ? Z_FINISH : Z_NO_FLUSH;
            const int rc = deflate(&zs, flush);
            if (rc == Z_STREAM_ERROR || (flush == Z_FINISH && rc != Z_STREAM_END)) die("writing output: gzip error");
        }
        if (parts.empty()) {
            if (deflate(&zs, Z_FINISH) != Z_STREAM_END) die("writing output: gzip error");
        }
        out.resize(zs.total_out);
        deflateEnd(&zs);
        return out;
    }

    void format_row(size_t i, std::string& b) const {
        const size_t n = ids->size();
        b.resize((*ids)[i].size() + n * (osep.size() + 11) + 16);
        char* p = &b[0];
        memcpy(p, (*ids)[i].data(), (*ids)[i].size());
        p += (*ids)[i].size();
        const uint16_t* r = mx->row(i);
        // kept columns [j0, j1): the rest print 0 (main.rs:245-253)
        const size_t j0 = layout == Layout::UpperTri ? i : 0, j1 = layout == Layout::LowerTri ? i + 1 : n;
        const uint32_t c = cap > 0 ? (uint32_t)cap : 0xffffffffu;
        if (cells.table) {
            const char* zero = &cells.slot[0];
            const size_t zl = cells.len[0];
            for (size_t j = 0; j < j0; j++, p += zl) memcpy(p, zero, 16);
            for (size_t j = j0; j < j1; j++) {
                const uint32_t v = std::min<uint32_t>(r[j], c);
                memcpy(p, &cells.slot[(size_t)v * 16], 16);
                p += cells.len[v];
            }
            for (size_t j = j1; j < n; j++, p += zl) memcpy(p, zero, 16);
        } else {
            for (size_t j = 0; j < n; j++) {
                memcpy(p, osep.data(), osep.size());
                p += osep.size();
                p = fmt_u(p, j >= j0 && j < j1 ? std::min<uint32_t>(r[j], c) : 0u);
            }
        }
        *p++ = '\n';
        b.resize((size_t)(p - &b[0]));
    }

    // Rows next .. i1-1 (all their stripes have arrived).
    void rows_until(size_t i1) {
        open();
        while (next < i1) {
            const size_t i0 = next, e = std::min(i1, i0 + 16 * GROUP);
            const size_t groups = (e - i0 + GROUP - 1) / GROUP;
            std::vector<std::vector<std::string>> text(groups);
            std::vector<std::string> packed(groups);
#pragma omp parallel for schedule(dynamic, 1)
            for (long g = 0; g < (long)groups; g++) {
                const size_t a = i0 + (size_t)g * GROUP, z = std::min(e, a + GROUP);
                text[g].resize(z - a);
                for (size_t i = a; i < z; i++) format_row(i, text[g][i - a]);
                if (gz) {
                    packed[g] = deflate_member(text[g]);
                    std::vector<std::string>().swap(text[g]);
                }
            }
            for (size_t g = 0; g < groups; g++) {
                if (gz) put(packed[g].data(), packed[g].size());
                else
                    for (auto& t : text[g]) put(t.data(), t.size());
            }
            next = e;
        }
    }

    void finish() {
        open();
        if (fflush(f) != 0) die("writing output: " + io_error(errno));
        if (f != stdout) fclose(f);
    }
};

struct StripeCtx {
    Matrix* mx;
    Writer* w;
};

// Stripe callback (tch_cb): rows r0.. of the stripe (columns j = r0 + c,
// c > r kept) into the matrix and mirrored into the later rows in 64 x 64
// blocks, then every row of the stripe is written while the device computes
// the next stripe.
void on_stripe(void* ctx, int r0, int rows, size_t w, const uint16_t* stripe) {
    StripeCtx* c = (StripeCtx*)ctx;
    Matrix& M = *c->mx;
    const size_t n = M.n, R0 = (size_t)r0;
#pragma omp parallel for schedule(static)
    for (int r = 0; r < rows; r++) {
        M.row(R0 + r)[R0 + r] = 0;
        memcpy(M.row(R0 + r) + R0 + r + 1, stripe + (size_t)r * w + r + 1, (w - r - 1) * 2);
    }
    const size_t jb = (n - R0 + 63) / 64;
#pragma omp parallel for schedule(dynamic, 4)
    for (long b = 0; b < (long)jb; b++) {
        const size_t j0 = R0 + (size_t)b * 64, j1 = std::min(n, j0 + 64);
        for (size_t i0 = R0; i0 < R0 + rows; i0 += 64) {
            const size_t i1 = std::min(R0 + rows, i0 + 64);
            for (size_t j = j0; j < j1; j++) {
                uint16_t* dst = M.row(j);
                for (size_t i = i0; i < i1 && i < j; i++) dst[i] = M.row(i)[j];
            }
        }
    }
    c->w->rows_until(R0 + rows);
}

void log_line(void* ctx, const char* line) {
    if (!*(bool*)ctx) fprintf(stderr, "%s\n", line);
}

}  // namespace

int main(int argc, char** argv) {
    Args args = parse_args(argc, argv);
    const double t_start = now();
    auto log = [&](const std::string& m) {
        if (!args.silent) fprintf(stderr, "%s\n", m.c_str());
    };
    char msg[512];
    if (args.threads > 0) omp_set_num_threads((int)args.threads);
    if (args.input_sep.size() != 1) die("input separator must be a single character");
    const char sep = args.input_sep[0];
    if ((size_t)std::max(tch_device_count(), 0) <= args.device)
        die("CUDA device " + std::to_string(args.device) + " not available");  // main.rs:283, kept verbatim

    // ---- load (main.rs:286-311) ----
    double t = now();
    std::string text;
    int eno = 0;
    if (!read_all(args.input, &text, &eno)) die("reading " + args.input + ": " + io_error(eno));
    if (!valid_utf8(text)) die("reading " + args.input + ": stream did not contain valid UTF-8");
    // Rust str::lines(): split on '\n', drop the '\r' of a "\r\n", no final
    // empty line after a trailing '\n'; a bare '\r' ending the text stays.
    std::vector<std::pair<size_t, size_t>> lines;
    for (size_t p = 0; p < text.size();) {
        size_t q = text.find('\n', p);
        const size_t end = q == std::string::npos ? text.size() : q;
        size_t e = end;
        if (q != std::string::npos && e > p && text[e - 1] == '\r') e--;
        lines.emplace_back(p, e);
        p = q == std::string::npos ? text.size() : q + 1;
    }
    if (lines.empty()) die("empty input");
    size_t l = 0;
    for (size_t p = lines[0].first; p < lines[0].second; p++) l += text[p] == sep;
    std::vector<std::pair<size_t, size_t>> data;
    for (size_t k = 1; k < lines.size(); k++)
        if (lines[k].second > lines[k].first) data.push_back(lines[k]);
    size_t n = data.size();
    std::vector<std::string> ids(n);
    std::vector<uint64_t> calls(n * l, 0);
    {
        std::vector<std::string> errs(n);
#pragma omp parallel for schedule(dynamic, 64)
        for (long ii = 0; ii < (long)n; ii++) {
            const size_t i = (size_t)ii;
            const char* b = text.data() + data[i].first;
            const char* e = text.data() + data[i].second;
            const char* f = (const char*)memchr(b, sep, (size_t)(e - b));
            const char* idend = f ? f : e;
            ids[i].assign(b, idend);
            if (!f) continue;
            const char* c = f + 1;
            size_t k = 0;
            while (true) {
                const char* ce = (const char*)memchr(c, sep, (size_t)(e - c));
                if (!ce) ce = e;
                uint64_t v;
                if (!parse_allele(c, ce, args.skip_repl, args.missing, &v)) {
                    const char* tb = c;
                    const char* te = ce;
                    trim(tb, te);
                    errs[i] = "sample " + ids[i] + ": allele '" + std::string(tb, te) + "' too large";
                    break;
                }
                if (k < l) calls[i * l + k] = v;
                k++;
                if (ce == e) break;
                c = ce + 1;
            }
        }
        for (auto& e : errs)
            if (!e.empty()) die(e);
    }
    snprintf(msg, sizeof msg, "Loaded %zu samples x %zu loci in %.2fs", n, l, now() - t);
    log(msg);

    // ---- completeness filters (main.rs:313-351) ----
    if (args.have_lc) {
        std::vector<size_t> keep;
        for (size_t k = 0; k < l; k++) {
            size_t c = 0;
            for (size_t i = 0; i < n; i++) c += calls[i * l + k] != 0;
            if (((double)c / (double)n) * 100.0 >= args.lc) keep.push_back(k);
        }
        if (keep.size() != l) {
            std::vector<uint64_t> nc(n * keep.size());
            for (size_t i = 0; i < n; i++)
                for (size_t d = 0; d < keep.size(); d++) nc[i * keep.size() + d] = calls[i * l + keep[d]];
            calls.swap(nc);
            l = keep.size();
        }
        snprintf(msg, sizeof msg, "Locus completeness >= %g%%: kept %zu loci", args.lc, l);
        log(msg);
    }
    if (args.have_sc) {
        std::vector<size_t> keep;
        for (size_t i = 0; i < n; i++) {
            size_t c = 0;
            for (size_t k = 0; k < l; k++) c += calls[i * l + k] != 0;
            if (((double)c / (double)l) * 100.0 >= args.sc) keep.push_back(i);
        }
        if (keep.size() != n) {
            std::vector<uint64_t> nc(keep.size() * l);
            std::vector<std::string> nids;
            for (size_t d = 0; d < keep.size(); d++) {
                std::copy(calls.begin() + keep[d] * l, calls.begin() + (keep[d] + 1) * l, nc.begin() + d * l);
                nids.push_back(ids[keep[d]]);
            }
            calls.swap(nc);
            ids.swap(nids);
            n = keep.size();
        }
        snprintf(msg, sizeof msg, "Sample completeness >= %g%%: kept %zu samples", args.sc, n);
        log(msg);
    }
    if (l > 65535) die("more than 65535 loci (u16 distances)");

    // ---- output (main.rs:472-490): written as the stripes arrive ----
    Writer wr;
    wr.ids = &ids;
    wr.osep = args.output_sep;
    wr.index_name = args.index_name;
    wr.layout = args.layout;
    wr.cap = args.max_dist;
    wr.to_file = args.have_output;
    wr.path = args.output;
    wr.gz = args.gzip || (args.have_output && args.output.size() >= 3 &&
                          args.output.compare(args.output.size() - 3, 3, ".gz") == 0);

    // ---- encode (main.rs:353-445), on the device ----
    t = now();
    char err[512] = {0};
    Matrix mx;
    {  // the full n x n mirror (7.9 GB at 63k) must fit in host memory
        const double need = (double)n * (double)n * 2.0, avail = host_available_bytes();
        if (avail > 0 && need > avail) {
            snprintf(msg, sizeof msg, "the %zu x %zu distance matrix needs %.2f GB of host memory, %.2f GB available",
                     n, n, need / 1e9, avail / 1e9);
            die(msg);
        }
    }
    mx.init(n);  // every cell is written before its row is formatted
    wr.mx = &mx;
    wr.cells.build(args.output_sep, (uint32_t)std::max<size_t>(l, 1));
    cgh_dataset* ds = NULL;
    const size_t stripe = std::max<size_t>(args.stripe, 8) / 8 * 8;
    // a device failure is reported the way main.rs:460 reports tch_run's
    auto device_error = [&]() { return strncmp(err, "HIP error", 9) ? std::string(err) : std::string("CUDA: ") + err; };
    if (n >= 2 && l > 0) {
        const size_t out_bytes = 3 * ((stripe * n + 7) / 8 * 8) * 2;  // the stripe ring's device slots
        const int tau_i = (int)std::min<size_t>(args.tau, (size_t)n + 1);  // counts never exceed n; 0 = auto
        int tau_used = 0;
        if ((uint64_t)n * l < (1ull << 31)) {
            if (cgh_dataset_encode((int)args.device, calls.data(), (int)n, (int)l, tau_i, out_bytes, log_line,
                                   &args.silent, &ds, &tau_used, NULL, err, sizeof err))
                die(device_error());
        } else {  // beyond one device sort: ranks and rare groups on the host, O/P built on the device
            std::vector<uint16_t> X(n * l);
            if (cgh_dense_ranks(calls.data(), (int)n, (int)l, X.data(), err, sizeof err)) die(err);
            tau_used = tau_i;
            if (!tau_used) {
                std::vector<uint32_t> cnt(cgh_allele_counts(X.data(), (int)n, (int)l, NULL, 0) + 1);
                const size_t nc = cgh_allele_counts(X.data(), (int)n, (int)l, cnt.data(), cnt.size());
                const size_t free_b = tch_free_vram((int)args.device);
                tau_used = cgh_choose_tau(cnt.data(), nc, (int)n, (int)l, out_bytes, free_b, log_line, &args.silent);
                if (!tau_used) {
                    snprintf(msg, sizeof msg, "no tau fits in %.2f GB of free VRAM", free_b / 1e9);
                    die(msg);
                }
            }
            if (cgh_dataset_build((int)args.device, X.data(), (int)n, (int)l, tau_used, &ds, err, sizeof err))
                die(device_error());
        }
        size_t D = 0;
        uint64_t n_rare = 0;
        cgh_dataset_info(ds, NULL, NULL, NULL, &D, &n_rare);
        if (args.max_dist > 0) cgh_dataset_set_cap(ds, args.max_dist);  // -X on the device
        snprintf(msg, sizeof msg, "Encoded in %.2fs: tau=%d, one-hot cols=%zu, rare equal pairs=%llu", now() - t,
                 tau_used, D, (unsigned long long)n_rare);
        log(msg);
    } else {  // nothing to compute; the rank limit still applies
        std::vector<uint16_t> X(n * l);
        if (cgh_dense_ranks(calls.data(), (int)n, (int)l, X.data(), err, sizeof err)) die(err);
    }
    std::vector<uint64_t>().swap(calls);
    if (!ds) mx.zero();  // nothing computed: an all-zero matrix (main.rs:466-470)
    if (ds) {
        // ---- GPU (main.rs:447-465) ----
        t = now();
        double gpu_secs = 0;
        StripeCtx sc = {&mx, &wr};
        const int rc = cgh_dataset_stream(ds, (int)stripe, on_stripe, &sc, &gpu_secs, err, sizeof err);
        cgh_dataset_free(ds);
        if (rc) die(std::string("CUDA: ") + err);  // main.rs:460
        snprintf(msg, sizeof msg, "Computed %zu x %zu matrix on MFMA in %.2fs (GPU pipeline %.2fs, rare pairs on GPU)",
                 n, n, now() - t, gpu_secs);
        log(msg);
    }

    // ---- the rows not written yet (all of them when nothing ran on the device) ----
    t = now();
    wr.rows_until(n);
    wr.finish();
    snprintf(msg, sizeof msg, "Wrote output in %.2fs. Total %.2fs", now() - t, now() - t_start);
    log(msg);
    return 0;
}
