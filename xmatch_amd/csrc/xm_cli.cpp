// xm_cli.cpp -- command-line crossmatch of two record files on one MI355X.
//
// Mirrors the reference's executable (BoostXmatch/BoostXmatch/BoostXmatch.cpp:
// 95-330): the same flags, defaults and exit codes, the same input format (raw
// 24-byte little-endian {int64 id; f64 ra_deg; f64 dec_deg}, count = file bytes /
// 24, BoostXmatch.cpp:251,281) and the same unit of work, a job of one A segment
// and one B segment (JobManager.cpp:5-13).
//
// Work split (the reference cut both files into id-order segments of -n objects,
// kept every A segment on the host and joined every A segment with every B
// segment, BoostXmatch.cpp:244-330):
//  * file A is cut into declination bands -- runs of whole zones -- of at most
//    -a rows (default: what a third of the free device memory holds, and at most
//    -n when given), from a zone histogram of A taken in one read of the file.  An
//    A that fits in one band is one band and needs no histogram;
//  * band k's A rows are resident on the GPU while file B is streamed past them:
//    only the B rows of the band's zones and nz + 2 zones either side (the zones
//    its rows can match, plus a margin) are kept, in rounds of at most -b rows
//    (default: from free device memory), cut into -n segments when given;
//  * each (band, round) is one batched join of the band with the round's
//    segments, written to <outdir>/<k>, k counting the (band, round) pairs in
//    order; every A row is in one band, so no pair is written twice.
// Differences, on purpose:
//  * the next round of B is read into page-locked memory by a second thread while
//    the GPU joins the current one;
//  * records are {idA, idB, dist_arcsec} (24 bytes) unless -f 16 asks for the
//    reference's {idA, idB} (Worker.cu:410-418);
//  * nothing is truncated (-m only sizes the first result buffer, per job);
//  * without -n a band is one segment (the reference left num_obj uninitialised);
//  * <outdir>/manifest records the inputs, the options and the band and round
//    plan; -c resumes with that plan and refuses other inputs or options.
#include <fcntl.h>
#include <ftw.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cctype>
#include <cerrno>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>
#include <thread>
#include <vector>

#include "../../include/xmatch.h"

namespace {

struct Options {
    std::string outdir;          // empty: count only, write nothing
    uint64_t maxout = 0;         // -m: first result buffer per job (0 = library default)
    double sr_arcsec = 5.0;      // -r
    double zh_arcsec = 0.0;      // -z (0 = radius)
    uint32_t threads = 0;        // -t (GPUs; the library drives one)
    uint64_t num_obj = 0;        // -n (0 = whole file)
    uint32_t verbosity = 0;      // -v
    bool overwrite = false;      // -x
    int format = 24;             // -f: 24 or 16 bytes per pair
    bool fp32 = false;           // -p fp32: FP32 filter + FP64 recheck
    uint64_t round_rows = 0;     // -b: B rows per round (0 = from free memory)
    uint64_t band_rows = 0;      // -a: A rows per band (0 = from free memory and -n)
    bool resume = false;         // -c: keep the output directory, skip finished rounds
    std::vector<std::string> inputs;
};

const char* USAGE =
    "Options:\n"
    "  -o [ --outdir ] arg (=out)    output directory path\n"
    "  -m [ --maxout ] arg (=0)      first result buffer per job (0 = 2 max(nA, nB));\n"
    "                                results are never truncated\n"
    "  -r [ --radius ] arg (=5)      search radius in arcsec\n"
    "  -z [ --zoneheight ] arg (=0)  zone height in arcsec, defaults to radius\n"
    "  -t [ --threads ] arg (=0)     number of GPUs (the library drives one)\n"
    "  -n [ --numobj ] arg           number of objects per segment (default: whole file)\n"
    "  -v [ --verbose ] [=arg(=9)]   verbosity\n"
    "  -x [ --overwrite ]            overwrite output directory\n"
    "  -f [ --format ] arg (=24)     24: {idA, idB, dist_arcsec}; 16: {idA, idB}\n"
    "  -p [ --precision ] arg        fp64 (default) or fp32 (FP32 filter, FP64 recheck)\n"
    "  -b [ --roundrows ] arg (=0)   B records per round (0 = from free device memory)\n"
    "  -a [ --bandrows ] arg (=0)    A records per declination band (0 = from free device\n"
    "                                memory, at most -n)\n"
    "  -c [ --resume ]               keep the output directory and skip the rounds whose\n"
    "                                file is complete (same inputs and options)\n"
    "  -h [ --help ]                 print help message\n";

void usage(const char* prog) {
    std::cout << "Usage: " << prog << " [options] file(s)" << std::endl << USAGE;
}

bool parse_u64(const std::string& s, uint64_t* v) {
    if (s.empty() || !std::isdigit(static_cast<unsigned char>(s[0]))) return false;
    char* end = nullptr;
    errno = 0;
    const unsigned long long x = std::strtoull(s.c_str(), &end, 10);
    if (errno || *end) return false;
    *v = x;
    return true;
}

bool parse_double(const std::string& s, double* v) {
    if (s.empty()) return false;
    char* end = nullptr;
    errno = 0;
    const double x = std::strtod(s.c_str(), &end);
    if (errno || *end) return false;
    *v = x;
    return true;
}

uint64_t file_size(const std::string& path) {
    struct stat st;
    if (stat(path.c_str(), &st) != 0 || !S_ISREG(st.st_mode)) return 0;
    return uint64_t(st.st_size);
}

// Returns 0, or the reference's codes: 1 bad option / help, 2 not two inputs,
// 3 an input missing or empty (BoostXmatch.cpp:136-167).
int parse(int argc, char** argv, Options& o) {
    std::vector<std::string> a(argv + 1, argv + argc);
    bool have_out = false;
    for (size_t i = 0; i < a.size(); i++) {
        std::string t = a[i], val;
        bool has_val = false;
        if (t.size() > 2 && t[0] == '-' && t[1] == '-') {
            const size_t eq = t.find('=');
            if (eq != std::string::npos) {
                val = t.substr(eq + 1);
                has_val = true;
                t = t.substr(0, eq);
            }
            static const char* longs[][2] = {{"--outdir", "-o"},   {"--maxout", "-m"},
                                             {"--radius", "-r"},   {"--zoneheight", "-z"},
                                             {"--threads", "-t"},  {"--numobj", "-n"},
                                             {"--verbose", "-v"},  {"--overwrite", "-x"},
                                             {"--format", "-f"},   {"--precision", "-p"},
                                             {"--roundrows", "-b"}, {"--resume", "-c"},
                                             {"--bandrows", "-a"},
                                             {"--help", "-h"}};
            bool found = false;
            for (auto& m : longs)
                if (t == m[0]) {
                    t = m[1];
                    found = true;
                }
            if (!found) {
                usage(argv[0]);
                std::cout << "Error: " << std::endl << "   unrecognised option '" << a[i] << "'"
                          << std::endl;
                return 1;
            }
        } else if (t.size() > 2 && t[0] == '-' && !std::isdigit(static_cast<unsigned char>(t[1]))) {
            val = t.substr(2);  // -oDIR, -r2.5
            has_val = true;
            t = t.substr(0, 2);
        }
        if (t.size() != 2 || t[0] != '-') {
            o.inputs.push_back(a[i]);
            continue;
        }
        const char c = t[1];
        // options whose value is optional: -o (out), -v (9)
        auto next_value = [&](bool optional, std::string* out) -> bool {
            if (has_val) {
                *out = val;
                return true;
            }
            if (i + 1 < a.size() && (a[i + 1].empty() || a[i + 1][0] != '-')) {
                if (optional && c == 'v') {
                    uint64_t tmp;
                    if (!parse_u64(a[i + 1], &tmp)) return false;  // a file name follows
                }
                *out = a[++i];
                return true;
            }
            return false;
        };
        std::string v;
        bool ok = true;
        switch (c) {
            case 'h': usage(argv[0]); return 1;
            case 'x': o.overwrite = true; break;
            case 'c': o.resume = true; break;
            case 'o':
                o.outdir = next_value(true, &v) ? v : "out";
                have_out = true;
                break;
            case 'v': {
                uint64_t lv = 9;
                if (next_value(true, &v)) ok = parse_u64(v, &lv);
                o.verbosity = uint32_t(lv);
                break;
            }
            case 'm': ok = next_value(false, &v) && parse_u64(v, &o.maxout); break;
            case 'r': ok = next_value(false, &v) && parse_double(v, &o.sr_arcsec); break;
            case 'z': ok = next_value(false, &v) && parse_double(v, &o.zh_arcsec); break;
            case 'n': ok = next_value(false, &v) && parse_u64(v, &o.num_obj); break;
            case 'b': ok = next_value(false, &v) && parse_u64(v, &o.round_rows); break;
            case 'a': ok = next_value(false, &v) && parse_u64(v, &o.band_rows); break;
            case 't': {
                uint64_t t64 = 0;
                ok = next_value(false, &v) && parse_u64(v, &t64);
                o.threads = uint32_t(t64);
                break;
            }
            case 'f': {
                uint64_t f = 0;
                ok = next_value(false, &v) && parse_u64(v, &f) && (f == 16 || f == 24);
                o.format = int(f);
                break;
            }
            case 'p':
                ok = next_value(false, &v) && (v == "fp64" || v == "fp32");
                o.fp32 = v == "fp32";
                break;
            default:
                ok = false;
        }
        if (!ok) {
            usage(argv[0]);
            std::cout << "Error: " << std::endl << "   bad value for option '" << a[i] << "'"
                      << std::endl;
            return 1;
        }
    }
    (void)have_out;
    if (!(o.sr_arcsec > 0.0) || o.zh_arcsec < 0.0) {
        usage(argv[0]);
        std::cout << "Error: " << std::endl << "   radius must be > 0, zone height >= 0" << std::endl;
        return 1;
    }
    if (o.inputs.size() != 2) {
        usage(argv[0]);
        std::cout << "Error: " << std::endl << "   Input 2 files!" << std::endl;
        std::cout << "Got this: " << std::endl;
        for (auto& f : o.inputs) std::cout << f << std::endl;
        return 2;
    }
    if (o.zh_arcsec == 0.0) o.zh_arcsec = o.sr_arcsec;  // BoostXmatch.cpp:158
    if (file_size(o.inputs[0]) < sizeof(xm_obj) || file_size(o.inputs[1]) < sizeof(xm_obj)) {
        usage(argv[0]);
        std::cout << "Error: " << std::endl << "   File not found! " << std::endl;
        return 3;
    }
    return 0;
}

bool is_dir(const std::string& p) {
    struct stat st;
    return stat(p.c_str(), &st) == 0 && S_ISDIR(st.st_mode);
}

int remove_entry(const char* path, const struct stat*, int, struct FTW*) {
    return std::remove(path);
}

bool remove_tree(const std::string& p) {
    return nftw(p.c_str(), remove_entry, 16, FTW_DEPTH | FTW_PHYS) == 0;
}

// Output directory handling of BoostXmatch.cpp:169-193: create; if it exists,
// replace it with -x, else ask.  Returns false to stop (exit -1).
bool prepare_outdir(const Options& o) {
    if (o.outdir.empty()) return true;
    if (!is_dir(o.outdir)) return mkdir(o.outdir.c_str(), 0777) == 0;
    if (o.resume) return true;
    if (!o.overwrite) {
        std::cout << "Output directory already exists: " << o.outdir << std::endl
                  << "Would like to overwrite this directory?[yes/no] " << std::endl;
        std::string resp;
        std::cin >> resp;
        for (auto& ch : resp) ch = char(std::toupper(static_cast<unsigned char>(ch)));
        if (resp != "YES") {
            std::cout << "use option -x to overwrite the current output directory" << std::endl;
            return false;
        }
    } else if (o.verbosity) {
        std::cerr << "Overwrite option set. Removing old directory." << std::endl;
    }
    return remove_tree(o.outdir) && mkdir(o.outdir.c_str(), 0777) == 0;
}

// A failure that ends the run with exit code `code`; main prints the message.
struct CliError {
    int code;
    std::string what;
};

struct Ctx {
    xm_ctx* c = nullptr;
    void check(xm_status st, const char* what) const {
        if (st != XM_OK)
            throw CliError{4, std::string(what) + ": " + xm_last_error_message(c)};
    }
};

// Reads records [first, first + n) of a file into dst (pread, any length).
bool read_span(int fd, uint64_t off, uint64_t bytes, char* p) {
    while (bytes) {
        const ssize_t got = pread(fd, p, std::min<uint64_t>(bytes, uint64_t(1) << 30), off_t(off));
        if (got <= 0) return false;
        p += got;
        off += uint64_t(got);
        bytes -= uint64_t(got);
    }
    return true;
}

bool write_span(int fd, uint64_t off, uint64_t bytes, const char* p) {
    while (bytes) {
        const ssize_t put = pwrite(fd, p, std::min<uint64_t>(bytes, uint64_t(1) << 30), off_t(off));
        if (put <= 0) return false;
        p += put;
        off += uint64_t(put);
        bytes -= uint64_t(put);
    }
    return true;
}

// Records [first, first + n) of the file, read by up to READ_THREADS threads in
// parallel spans (one thread copies from the page cache at a few GB/s).
constexpr unsigned READ_THREADS = 8;
bool read_records(int fd, uint64_t first, uint64_t n, xm_obj* dst) {
    const uint64_t bytes = n * sizeof(xm_obj), off = first * sizeof(xm_obj);
    char* p = reinterpret_cast<char*>(dst);
    const uint64_t min_span = uint64_t(64) << 20;
    const unsigned k = unsigned(std::min<uint64_t>(READ_THREADS, std::max<uint64_t>(1, bytes / min_span)));
    if (k <= 1) return read_span(fd, off, bytes, p);
    const uint64_t span = (bytes + k - 1) / k;
    std::vector<std::thread> th;
    std::vector<char> ok(k, 0);
    for (unsigned t = 0; t < k; t++) {
        const uint64_t a = std::min<uint64_t>(bytes, t * span), b = std::min<uint64_t>(bytes, a + span);
        th.emplace_back([&, t, a, b] { ok[t] = read_span(fd, off + a, b - a, p + a); });
    }
    for (auto& x : th) x.join();
    for (char c : ok)
        if (!c) return false;
    return true;
}

// Round r's pairs to <path>: copied back into page-locked memory, then formatted
// and written by up to WRITE_THREADS threads, each into its own slice of the file
// (records are fixed-size), under a temporary name that is renamed when the file
// is complete (what -c relies on).
constexpr unsigned WRITE_THREADS = 8;
bool write_round(const Ctx& C, const xm_result* res, uint64_t m, const std::string& path,
                 int format) {
    int64_t* ia = nullptr;
    int64_t* ib = nullptr;
    double* dist = nullptr;
    const uint64_t n_alloc = std::max<uint64_t>(m, 1);
    struct Free {
        const Ctx& C;
        int64_t** ia;
        int64_t** ib;
        double** dist;
        ~Free() {
            xm_host_free(C.c, *ia);
            xm_host_free(C.c, *ib);
            xm_host_free(C.c, *dist);
        }
    } free_all{C, &ia, &ib, &dist};
    C.check(xm_host_alloc(C.c, n_alloc * 8, reinterpret_cast<void**>(&ia)), "pinned output");
    C.check(xm_host_alloc(C.c, n_alloc * 8, reinterpret_cast<void**>(&ib)), "pinned output");
    C.check(xm_host_alloc(C.c, n_alloc * 8, reinterpret_cast<void**>(&dist)), "pinned output");
    C.check(xm_result_copy(C.c, res, ia, ib, dist), "copy");
    const std::string part = path + ".part";
    const int fd = open(part.c_str(), O_WRONLY | O_CREAT | O_TRUNC, 0666);
    bool ok = fd >= 0;
    if (!ok) std::cerr << "xmatch: cannot write " << part << std::endl;
    const size_t rec = size_t(format);
    if (ok && m) {
        const uint64_t min_rows = uint64_t(1) << 20;
        const unsigned k = unsigned(std::min<uint64_t>(WRITE_THREADS, std::max<uint64_t>(1, m / min_rows)));
        const uint64_t per = (m + k - 1) / k;
        std::vector<std::thread> th;
        std::vector<char> good(k, 1);
        for (unsigned t = 0; t < k; t++) {
            th.emplace_back([&, t] {
                const uint64_t a = std::min<uint64_t>(m, t * per), b = std::min<uint64_t>(m, a + per);
                const uint64_t chunk = uint64_t(1) << 18;
                std::vector<char> buf(size_t(std::min<uint64_t>(b - a, chunk)) * rec + 1);
                for (uint64_t k0 = a; k0 < b; k0 += chunk) {
                    const uint64_t k1 = std::min<uint64_t>(b, k0 + chunk);
                    char* q = buf.data();
                    for (uint64_t i = k0; i < k1; i++, q += rec) {
                        std::memcpy(q, &ia[i], 8);
                        std::memcpy(q + 8, &ib[i], 8);
                        if (rec == 24) std::memcpy(q + 16, &dist[i], 8);
                    }
                    if (!write_span(fd, k0 * rec, (k1 - k0) * rec, buf.data())) {
                        good[t] = 0;
                        return;
                    }
                }
            });
        }
        for (auto& x : th) x.join();
        for (char g : good) ok = ok && g;
        if (!ok) std::cerr << "xmatch: write error on " << path << std::endl;
    }
    if (fd >= 0 && close(fd) != 0) ok = false;
    if (ok && std::rename(part.c_str(), path.c_str()) != 0) {
        std::cerr << "xmatch: cannot finish " << path << std::endl;
        ok = false;
    }
    return ok;
}

// Seconds since t0, for the -v 2 phase times.
double since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// Segments of at most num_obj rows over [0, n) of a device record array.
std::vector<xm_segment> cut(const xm_obj* d, uint64_t n, uint64_t num_obj, double h_deg) {
    std::vector<xm_segment> s;
    const uint64_t step = num_obj ? num_obj : std::max<uint64_t>(n, 1);
    for (uint64_t at = 0; at < n; at += step) s.push_back({d + at, std::min(step, n - at), h_deg});
    return s;
}

// The zone of a declination as the library assigns it (Sorter.cu:20-24, with
// dec = +90 clamped into the last zone).
int64_t zone_of(double dec, double h_deg, int64_t n_zones) {
    const int64_t z = int64_t(std::floor((dec + 90.0) / h_deg));
    return z < 0 ? 0 : (z > n_zones - 1 ? n_zones - 1 : z);
}

// A record file read front to back in chunks, keeping the records of the zones
// [z_lo, z_hi]: each next() appends up to cap of them to dst, continuing where
// the last call stopped.  A range that covers every zone reads straight into dst.
constexpr uint64_t READ_CHUNK = uint64_t(1) << 22;  // records per read (96 MB)
class ZoneReader {
  public:
    ZoneReader(int fd, uint64_t n, double h_deg, int64_t n_zones, int64_t z_lo, int64_t z_hi)
        : fd_(fd), n_(n), h_(h_deg), nz_(n_zones), lo_(std::max<int64_t>(z_lo, 0)),
          hi_(std::min<int64_t>(z_hi, n_zones - 1)), all_(lo_ == 0 && hi_ == n_zones - 1) {}
    // false on a read error; *got = rows appended (0 at the end of the file)
    bool next(xm_obj* dst, uint64_t cap, uint64_t* got) {
        *got = 0;
        if (all_) {
            const uint64_t take = std::min(cap, n_ - next_rec_);
            if (take && !read_records(fd_, next_rec_, take, dst)) return false;
            next_rec_ += take;
            *got = take;
            return true;
        }
        while (*got < cap) {
            if (pos_ == len_) {
                if (next_rec_ == n_) break;
                len_ = std::min(READ_CHUNK, n_ - next_rec_);
                chunk_.resize(len_);
                if (!read_records(fd_, next_rec_, len_, chunk_.data())) return false;
                next_rec_ += len_;
                pos_ = 0;
            }
            for (; pos_ < len_ && *got < cap; pos_++) {
                const int64_t z = zone_of(chunk_[pos_].dec_deg, h_, nz_);
                if (z >= lo_ && z <= hi_) dst[(*got)++] = chunk_[pos_];
            }
        }
        return true;
    }

  private:
    int fd_;
    uint64_t n_, next_rec_ = 0;
    double h_;
    int64_t nz_, lo_, hi_;
    bool all_;
    std::vector<xm_obj> chunk_;
    uint64_t pos_ = 0, len_ = 0;
};

// The run's work plan: A's declination bands (zone ranges, in order) with their
// row counts, and B's rows per round.
struct Plan {
    uint64_t round_rows = 0;
    std::vector<int64_t> lo, hi;
    std::vector<uint64_t> a_rows;
};

// Bands of whole zones holding at most `budget` A rows each (a zone bin holding
// more is a band of its own), from a histogram of A's zones in bins of w zones,
// w = ceil(n_zones / 2^20); bands without A rows are left out.
Plan plan_bands(int fa, uint64_t nA, double h_deg, int64_t n_zones, uint64_t budget) {
    Plan p;
    if (nA <= budget) {
        p.lo.push_back(0);
        p.hi.push_back(n_zones - 1);
        p.a_rows.push_back(nA);
        return p;
    }
    const int64_t w = (n_zones + (int64_t(1) << 20) - 1) >> 20;
    const int64_t nbins = (n_zones + w - 1) / w;
    std::vector<uint64_t> hist(size_t(nbins), 0);
    std::vector<xm_obj> chunk;
    for (uint64_t at = 0; at < nA; at += READ_CHUNK) {
        const uint64_t len = std::min(READ_CHUNK, nA - at);
        chunk.resize(len);
        if (!read_records(fa, at, len, chunk.data())) throw CliError{3, "read error on file A"};
        for (const xm_obj& o : chunk) hist[size_t(zone_of(o.dec_deg, h_deg, n_zones) / w)]++;
    }
    int64_t start = 0;
    uint64_t acc = 0;
    auto close = [&](int64_t end_bin) {  // bins [start, end_bin)
        if (acc) {
            p.lo.push_back(start * w);
            p.hi.push_back(std::min(end_bin * w, n_zones) - 1);
            p.a_rows.push_back(acc);
        }
        start = end_bin;
        acc = 0;
    };
    for (int64_t b = 0; b < nbins; b++) {
        if (acc && acc + hist[size_t(b)] > budget) close(b);
        acc += hist[size_t(b)];
    }
    close(nbins);
    return p;
}

// ---- the manifest: inputs, options and plan of the run that wrote <outdir>
std::string file_identity(const std::string& path) {
    struct stat st;
    if (stat(path.c_str(), &st) != 0) return "missing " + path;
    return std::to_string(uint64_t(st.st_size)) + " " + std::to_string(int64_t(st.st_mtim.tv_sec)) +
           "." + std::to_string(int64_t(st.st_mtim.tv_nsec)) + " " + path;
}

std::string identity_lines(const Options& o, uint64_t num_obj_given) {
    char opt[256];
    std::snprintf(opt, sizeof(opt), "options r=%.17g z=%.17g n=%llu f=%d p=%s a=%llu b=%llu",
                  o.sr_arcsec, o.zh_arcsec, (unsigned long long)num_obj_given, o.format,
                  o.fp32 ? "fp32" : "fp64", (unsigned long long)o.band_rows,
                  (unsigned long long)o.round_rows);
    return "xmatch-manifest 1\na " + file_identity(o.inputs[0]) + "\nb " +
           file_identity(o.inputs[1]) + "\n" + opt + "\n";
}

std::string manifest_text(const std::string& identity, const Plan& p) {
    std::string t = identity + "plan round_rows=" + std::to_string(p.round_rows) +
                    " bands=" + std::to_string(p.lo.size()) + "\n";
    for (size_t k = 0; k < p.lo.size(); k++)
        t += "band " + std::to_string(p.lo[k]) + " " + std::to_string(p.hi[k]) + " " +
             std::to_string(p.a_rows[k]) + "\n";
    return t;
}

// The plan of a manifest whose identity lines equal `identity`; CliError 1 when
// the manifest is missing, unreadable or was written for other inputs or options.
Plan read_manifest(const std::string& path, const std::string& identity) {
    FILE* f = std::fopen(path.c_str(), "r");
    if (!f) throw CliError{1, "-c: no manifest in the output directory to resume from"};
    std::string text;
    char buf[4096];
    size_t got;
    while ((got = std::fread(buf, 1, sizeof(buf), f)) > 0) text.append(buf, got);
    std::fclose(f);
    if (text.compare(0, identity.size(), identity) != 0)
        throw CliError{1, "-c: the output directory was written with other inputs or options "
                          "(see its manifest); run without -c, or with the same ones"};
    Plan p;
    unsigned long long rr = 0, nb = 0;
    const char* q = text.c_str() + identity.size();
    int used = 0;
    if (std::sscanf(q, "plan round_rows=%llu bands=%llu\n%n", &rr, &nb, &used) != 2 || !rr)
        throw CliError{1, "-c: the manifest's plan is unreadable"};
    q += used;
    p.round_rows = rr;
    for (unsigned long long k = 0; k < nb; k++) {
        long long lo = 0, hi = 0;
        unsigned long long rows = 0;
        if (std::sscanf(q, "band %lld %lld %llu\n%n", &lo, &hi, &rows, &used) != 3)
            throw CliError{1, "-c: the manifest's plan is unreadable"};
        q += used;
        p.lo.push_back(lo);
        p.hi.push_back(hi);
        p.a_rows.push_back(rows);
    }
    return p;
}

void write_manifest(const std::string& path, const std::string& text) {
    const std::string part = path + ".part";
    FILE* f = std::fopen(part.c_str(), "w");
    bool ok = f && std::fwrite(text.data(), 1, text.size(), f) == text.size();
    if (f && std::fclose(f) != 0) ok = false;
    if (!ok || std::rename(part.c_str(), path.c_str()) != 0)
        throw CliError{4, "cannot write " + path};
}

// ---- owners that release on every way out of run()
struct CtxOwner {
    xm_ctx* c = nullptr;
    ~CtxOwner() { xm_ctx_destroy(c); }
};
struct FdOwner {
    int fd = -1;
    ~FdOwner() {
        if (fd >= 0) close(fd);
    }
};
struct HostOwner {
    const Ctx& C;
    xm_obj* p = nullptr;
    ~HostOwner() { xm_host_free(C.c, p); }
};
struct DevOwner {
    const Ctx& C;
    void* p = nullptr;
    ~DevOwner() {
        if (p) xm_device_free(C.c, p);
    }
};
struct CatalogOwner {
    xm_catalog* c = nullptr;
    ~CatalogOwner() { xm_catalog_destroy(c); }
};
struct ResultOwner {
    xm_result* r = nullptr;
    ~ResultOwner() { xm_result_destroy(r); }
};
struct ThreadOwner {
    std::thread t;
    ~ThreadOwner() {
        if (t.joinable()) t.join();
    }
};

// Device records of `n` host rows, prepared as a many-segment catalog of -n segments.
xm_catalog* prepare(const Ctx& C, const xm_obj* h, uint64_t n, uint64_t num_obj, double h_deg,
                    const char* what) {
    DevOwner d{C};
    C.check(xm_device_alloc(C.c, std::max<uint64_t>(n, 1) * sizeof(xm_obj), &d.p), what);
    C.check(xm_memcpy_h2d(C.c, d.p, h, n * sizeof(xm_obj)), what);
    auto segs = cut(static_cast<xm_obj*>(d.p), n, num_obj, h_deg);
    if (segs.size() > 65536)
        throw CliError{1, std::string(what) + ": more than 65536 segments; raise -n"};
    xm_catalog* cat = nullptr;
    C.check(xm_catalog_create_segments_device(C.c, segs.data(), segs.size(), &cat), what);
    return cat;
}

int run(Options& o) {
    const uint64_t num_obj_given = o.num_obj;
    const double sr_deg = o.sr_arcsec / 3600.0, zh_deg = o.zh_arcsec / 3600.0;
    const uint64_t nA = file_size(o.inputs[0]) / sizeof(xm_obj);
    const uint64_t nB = file_size(o.inputs[1]) / sizeof(xm_obj);
    const int64_t n_zones = int64_t(std::ceil(180.0 / zh_deg));
    const int64_t nz = int64_t(std::ceil(sr_deg / zh_deg));  // Worker.cu:313
    if (o.verbosity) {
        std::cerr << "Input file A   :  " << o.inputs[0] << " (" << nA << " records)\n"
                  << "Input file B   :  " << o.inputs[1] << " (" << nB << " records)\n"
                  << "Output dir     :  " << (o.outdir.empty() ? "(none)" : o.outdir) << "\n"
                  << "Search radius  :  " << o.sr_arcsec << "\n"
                  << "Zone height    :  " << o.zh_arcsec << "\n"
                  << "# of obj/seg   :  " << (o.num_obj ? std::to_string(o.num_obj) : "whole band")
                  << std::endl;
        if (o.threads > 1) std::cerr << "-t " << o.threads << ": the library drives one GPU\n";
        if (nB < nA) std::cerr << "!! Larger file in memory?!" << std::endl;  // BoostXmatch.cpp:240
    }
    if (!prepare_outdir(o)) return 255;
    constexpr uint64_t MAX_ROWS = 4000000000ull;  // below the library's 2^32 - 8193 per catalog
    if (o.num_obj > MAX_ROWS) o.num_obj = MAX_ROWS;

    // a resumed run keeps the plan of the run that wrote the output directory, and
    // refuses other inputs or options before it touches the GPU
    Plan plan;
    const std::string manifest = o.outdir.empty() ? "" : o.outdir + "/manifest";
    const std::string identity = identity_lines(o, num_obj_given);
    const bool resumed = o.resume && !manifest.empty();
    if (resumed) plan = read_manifest(manifest, identity);

    const auto t_start = std::chrono::steady_clock::now();
    Ctx C;
    CtxOwner ctx_owner;
    if (xm_ctx_create(0, &C.c) != XM_OK)
        throw CliError{4, std::string("no usable GPU: ") + xm_last_error_message(nullptr)};
    ctx_owner.c = C.c;
    FdOwner fa{open(o.inputs[0].c_str(), O_RDONLY)}, fb{open(o.inputs[1].c_str(), O_RDONLY)};
    if (fa.fd < 0 || fb.fd < 0) throw CliError{3, "cannot open inputs"};

    // ---- a new plan: A's bands and B's rounds, sized from free device memory (a
    // band: records, sort words, rows and keys, ~100 B a row during its build; a
    // round: the same plus its pairs, ~200 B a row)
    if (!resumed) {
        uint64_t free_b = 0, total_b = 0;
        C.check(xm_ctx_mem_info(C.c, &free_b, &total_b), "memory info");
        uint64_t band = o.band_rows ? o.band_rows : std::max<uint64_t>(free_b / 3 / 100, 1);
        if (o.num_obj) band = std::min(band, o.num_obj);
        plan = plan_bands(fa.fd, nA, zh_deg, n_zones, std::min(band, MAX_ROWS));
        // a round never holds more than file B (its two page-locked buffers would
        // otherwise be sized for what the device could hold: 2 x 12 GB on one MI355X)
        plan.round_rows = o.round_rows ? o.round_rows
                                       : std::min(std::max<uint64_t>(free_b / 3 / 200, 1),
                                                  std::max<uint64_t>(nB, 1));
        plan.round_rows = std::min(plan.round_rows, MAX_ROWS);
        if (!manifest.empty()) write_manifest(manifest, manifest_text(identity, plan));
    }
    if (o.verbosity)
        std::cerr << "plan: " << plan.lo.size() << " declination band(s) of A, B in rounds of "
                  << plan.round_rows << " records" << std::endl;
    const double t_plan = since(t_start);

    uint64_t max_band = 1;
    for (uint64_t r : plan.a_rows) max_band = std::max(max_band, r);
    HostOwner ha{C}, hb0{C}, hb1{C};
    C.check(xm_host_alloc(C.c, max_band * sizeof(xm_obj), reinterpret_cast<void**>(&ha.p)),
            "page-locked A");
    C.check(xm_host_alloc(C.c, plan.round_rows * sizeof(xm_obj), reinterpret_cast<void**>(&hb0.p)),
            "page-locked B");
    C.check(xm_host_alloc(C.c, plan.round_rows * sizeof(xm_obj), reinterpret_cast<void**>(&hb1.p)),
            "page-locked B");
    xm_obj* hb[2] = {hb0.p, hb1.p};
    const double t_pin = since(t_start);
    if (o.verbosity >= 2)
        std::cerr << "times (s from start): plan " << t_plan << ", page-locked buffers " << t_pin
                  << std::endl;

    uint64_t total = 0, skipped = 0, k = 0;
    for (size_t band = 0; band < plan.lo.size(); band++) {
        // B's rows of the band's zones and nz + 2 zones either side: the first round
        // is read while A's band is read, uploaded and prepared
        ZoneReader rb(fb.fd, nB, zh_deg, n_zones, plan.lo[band] - nz - 2, plan.hi[band] + nz + 2);
        uint64_t got[2] = {0, 0};
        bool first_ok = true;
        ThreadOwner first;
        first.t = std::thread([&] { first_ok = rb.next(hb[0], plan.round_rows, &got[0]); });
        // A's rows of the band, resident while B streams past
        ZoneReader ra(fa.fd, nA, zh_deg, n_zones, plan.lo[band], plan.hi[band]);
        uint64_t na = 0;
        if (!ra.next(ha.p, plan.a_rows[band], &na) || na != plan.a_rows[band])
            throw CliError{3, "read error on " + o.inputs[0] + " (or it changed since the plan)"};
        const double t_read_a = since(t_start);
        CatalogOwner A{prepare(C, ha.p, na, o.num_obj, zh_deg, "prepare A")};
        uint64_t nsegA = 0;
        xm_catalog_segment_count(A.c, &nsegA);
        const double t_prep_a = since(t_start);
        first.t.join();
        if (!first_ok) throw CliError{3, "read error on " + o.inputs[1]};
        if (o.verbosity) {
            std::cerr << "band " << band << ": zones " << plan.lo[band] << ".." << plan.hi[band]
                      << ", " << na << " A records";
            if (o.verbosity >= 2)
                std::cerr << " (s from start: read " << t_read_a << ", prepared " << t_prep_a
                          << "; first B round read by " << since(t_start) << ")";
            std::cerr << std::endl;
        }
        for (int cur = 0; got[cur] > 0; cur ^= 1) {
            // the next round is read while this one joins
            bool next_ok = true;
            got[cur ^ 1] = 0;
            ThreadOwner reader;
            reader.t = std::thread([&, cur] {
                next_ok = rb.next(hb[cur ^ 1], plan.round_rows, &got[cur ^ 1]);
            });
            const std::string path = o.outdir.empty() ? "" : o.outdir + "/" + std::to_string(k);
            struct stat pst;
            if (o.resume && !path.empty() && stat(path.c_str(), &pst) == 0) {
                // finished by an earlier run: a round's file only appears complete
                if (o.verbosity) std::cerr << "round " << k << ": done before, skipped" << std::endl;
                skipped++;
            } else {
                const uint64_t n = got[cur];
                CatalogOwner B{prepare(C, hb[cur], n, o.num_obj, zh_deg, "prepare B")};
                uint64_t nsegB = 0;
                xm_catalog_segment_count(B.c, &nsegB);
                std::vector<xm_job> jobs;  // JobManager.cpp:5-13: every A segment with every B segment
                for (uint32_t i = 0; i < nsegA; i++)
                    for (uint32_t j = 0; j < nsegB; j++) jobs.push_back({i, j, sr_deg});
                ResultOwner res;
                xm_match_stats st;
                C.check(xm_match_batch(C.c, A.c, B.c, jobs.data(), jobs.size(),
                                       o.fp32 ? XM_FP32_FILTER_FP64_RECHECK : XM_FP64_PLAIN,
                                       o.maxout * jobs.size(), &res.r, &st),
                        "join");
                uint64_t m = 0;
                xm_result_count(res.r, &m);
                total += m;
                if (!o.outdir.empty() && !write_round(C, res.r, m, path, o.format))
                    throw CliError{4, "cannot write " + path};
                if (o.verbosity) {
                    std::cerr << "round " << k << ": band " << band << ", " << n << " B records, "
                              << jobs.size() << " jobs, " << m << " matches, join " << st.join_ms
                              << " ms";
                    if (o.verbosity >= 2) std::cerr << " (done " << since(t_start) << " s from start)";
                    std::cerr << std::endl;
                }
            }
            k++;
            reader.t.join();
            if (!next_ok) throw CliError{3, "read error on " + o.inputs[1]};
        }
    }
    const double secs = since(t_start);
    std::cout << total << " matches in " << secs << " s";
    if (skipped) std::cout << " (" << skipped << " round(s) done before)";
    std::cout << std::endl;
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    Options o;
    if (const int e = parse(argc, argv, o)) return e;
    try {
        return run(o);
    } catch (const CliError& e) {
        if (e.code == 1) std::cout << "Error: " << std::endl << "   " << e.what << std::endl;
        std::cerr << "xmatch: " << e.what << std::endl;
        return e.code;
    }
}
