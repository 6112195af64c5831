// xm_cli.cpp -- command-line crossmatch of two record files on one MI355X.
//
// Mirrors the reference's executable (BoostXmatch/BoostXmatch/BoostXmatch.cpp:
// 95-330): the same flags, defaults and exit codes, the same input format (raw
// 24-byte little-endian {int64 id; f64 ra_deg; f64 dec_deg}, count = file bytes /
// 24, BoostXmatch.cpp:251,281) and the same work split: file A is cut into
// segments of -n records and kept resident; file B is read in rounds of segments;
// each round joins every A segment with every B segment of the round
// (JobManager.cpp:5-13) and writes its pairs to <outdir>/<round>.
//
// Differences, on purpose:
//  * one batched join per round (xm_match_batch) instead of one Worker::Match per job;
//  * file A stays on the GPU (the reference kept it on the host and copied every
//    segment again per job);
//  * the next round of B is read into page-locked memory by a second thread while
//    the GPU joins the current round;
//  * records are {idA, idB, dist_arcsec} (24 bytes) unless -f 16 asks for the
//    reference's {idA, idB} (Worker.cu:410-418);
//  * nothing is truncated (-m only sizes the first result buffer, per job);
//  * without -n a file is one segment (the reference left num_obj uninitialised).
#include <fcntl.h>
#include <ftw.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cctype>
#include <cerrno>
#include <chrono>
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

struct Ctx {
    xm_ctx* c = nullptr;
    void check(xm_status st, const char* what) const {
        if (st != XM_OK) {
            std::cerr << "xmatch: " << what << ": " << xm_last_error_message(c) << std::endl;
            std::exit(4);
        }
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
    xm_host_free(C.c, ia);
    xm_host_free(C.c, ib);
    xm_host_free(C.c, dist);
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

}  // namespace

int main(int argc, char** argv) {
    Options o;
    if (const int e = parse(argc, argv, o)) return e;
    const double sr_deg = o.sr_arcsec / 3600.0, zh_deg = o.zh_arcsec / 3600.0;
    const uint64_t nA = file_size(o.inputs[0]) / sizeof(xm_obj);
    const uint64_t nB = file_size(o.inputs[1]) / sizeof(xm_obj);
    if (o.verbosity) {
        std::cerr << "Input file A   :  " << o.inputs[0] << " (" << nA << " records)\n"
                  << "Input file B   :  " << o.inputs[1] << " (" << nB << " records)\n"
                  << "Output dir     :  " << (o.outdir.empty() ? "(none)" : o.outdir) << "\n"
                  << "Search radius  :  " << o.sr_arcsec << "\n"
                  << "Zone height    :  " << o.zh_arcsec << "\n"
                  << "# of obj/seg   :  " << (o.num_obj ? std::to_string(o.num_obj) : "whole file")
                  << std::endl;
        if (o.threads > 1) std::cerr << "-t " << o.threads << ": the library drives one GPU\n";
        if (nB < nA) std::cerr << "!! Larger file in memory?!" << std::endl;  // BoostXmatch.cpp:240
    }
    if (!prepare_outdir(o)) return 255;
    if (o.num_obj >= 0xffffffffull) o.num_obj = 0xfffffffeull;
    if (!o.num_obj && (nA >= 0xffffffffull || nB >= 0xffffffffull)) o.num_obj = uint64_t(1) << 31;

    const auto t_start = std::chrono::steady_clock::now();
    Ctx C;
    if (xm_ctx_create(0, &C.c) != XM_OK) {
        std::cerr << "xmatch: no usable GPU: " << xm_last_error_message(nullptr) << std::endl;
        return 4;
    }
    const int fa = open(o.inputs[0].c_str(), O_RDONLY), fb = open(o.inputs[1].c_str(), O_RDONLY);
    if (fa < 0 || fb < 0) {
        std::cerr << "xmatch: cannot open inputs" << std::endl;
        return 3;
    }
    const double t_init = since(t_start);
    // ---- the rounds of file B: whole segments, the next round read ahead
    uint64_t round_rows = o.round_rows;
    if (!round_rows) {
        // a round needs its records, catalog build and output (~200 B per row);
        // take half of the free device memory, the context's cache counted free
        uint64_t free_b = 0, total_b = 0;
        C.check(xm_ctx_mem_info(C.c, &free_b, &total_b), "memory info");
        round_rows = std::max<uint64_t>(free_b / 2 / 200, 1);
    }
    const uint64_t seg_b = o.num_obj ? o.num_obj : std::max<uint64_t>(nB, 1);
    if (round_rows < seg_b) round_rows = seg_b;
    round_rows = std::min<uint64_t>(round_rows / seg_b * seg_b, 0xfffffffeull / seg_b * seg_b);
    std::vector<std::pair<uint64_t, uint64_t>> rounds;  // (first row, rows)
    for (uint64_t at = 0; at < nB; at += round_rows) rounds.push_back({at, std::min(round_rows, nB - at)});
    // Two page-locked buffers: file A goes through hb[0]; round r of B through
    // hb[(r + 1) & 1], so round 0 is read while A uploads and is prepared, and each
    // later round while the one before it joins.
    const uint64_t b_rows = rounds.empty() ? 1 : rounds[0].second;
    xm_obj* hb[2] = {nullptr, nullptr};
    C.check(xm_host_alloc(C.c, std::max(nA, b_rows) * sizeof(xm_obj), reinterpret_cast<void**>(&hb[0])),
            "pinned A");
    C.check(xm_host_alloc(C.c, b_rows * sizeof(xm_obj), reinterpret_cast<void**>(&hb[1])), "pinned B");
    const double t_pin = since(t_start);
    bool read_ok = true;
    std::thread b0_reader;
    if (!rounds.empty())
        b0_reader = std::thread([&] { read_ok = read_records(fb, rounds[0].first, rounds[0].second, hb[1]); });
    // ---- file A: resident on the device as one many-segment catalog
    xm_catalog* A = nullptr;
    double t_read_a = 0, t_prep_a = 0;
    {
        void* d = nullptr;
        if (!read_records(fa, 0, nA, hb[0])) {
            std::cerr << "xmatch: read error on " << o.inputs[0] << std::endl;
            if (b0_reader.joinable()) b0_reader.join();
            return 3;
        }
        t_read_a = since(t_start);
        C.check(xm_device_alloc(C.c, nA * sizeof(xm_obj), &d), "device A");
        C.check(xm_memcpy_h2d(C.c, d, hb[0], nA * sizeof(xm_obj)), "upload A");
        auto segs = cut(static_cast<xm_obj*>(d), nA, o.num_obj, zh_deg);
        if (segs.size() > 65536) {
            std::cerr << "xmatch: file A makes more than 65536 segments; raise -n" << std::endl;
            if (b0_reader.joinable()) b0_reader.join();
            return 1;
        }
        C.check(xm_catalog_create_segments_device(C.c, segs.data(), segs.size(), &A), "prepare A");
        C.check(xm_device_free(C.c, d), "device A");
        t_prep_a = since(t_start);
        if (o.verbosity) std::cerr << "A: " << segs.size() << " segment(s) prepared" << std::endl;
    }
    uint64_t nsegA = 0;
    xm_catalog_segment_count(A, &nsegA);
    if (b0_reader.joinable()) b0_reader.join();
    const double t_read_b0 = since(t_start);
    if (o.verbosity >= 2)
        std::cerr << "times (s from start): context " << t_init << ", page-locked buffers " << t_pin
                  << ", A read " << t_read_a << ", A prepared " << t_prep_a
                  << ", B round 0 read " << t_read_b0 << std::endl;
    uint64_t total = 0, skipped = 0;
    for (size_t r = 0; r < rounds.size() && read_ok; r++) {
        bool next_ok = true;
        std::thread reader;
        const std::string path = o.outdir.empty() ? "" : o.outdir + "/" + std::to_string(r);
        struct stat pst;
        if (o.resume && !path.empty() && stat(path.c_str(), &pst) == 0) {
            // finished by an earlier run: a round's file only appears complete
            if (o.verbosity) std::cerr << "round " << r << ": done before, skipped" << std::endl;
            skipped++;
            if (r + 1 < rounds.size())
                read_ok = read_records(fb, rounds[r + 1].first, rounds[r + 1].second, hb[r & 1]);
            continue;
        }
        if (r + 1 < rounds.size())
            reader = std::thread([&, r] {
                next_ok = read_records(fb, rounds[r + 1].first, rounds[r + 1].second, hb[r & 1]);
            });
        const uint64_t n = rounds[r].second;
        void* d = nullptr;
        C.check(xm_device_alloc(C.c, n * sizeof(xm_obj), &d), "device B");
        C.check(xm_memcpy_h2d(C.c, d, hb[(r + 1) & 1], n * sizeof(xm_obj)), "upload B");
        auto segs = cut(static_cast<xm_obj*>(d), n, o.num_obj, zh_deg);
        xm_catalog* B = nullptr;
        C.check(xm_catalog_create_segments_device(C.c, segs.data(), segs.size(), &B), "prepare B");
        C.check(xm_device_free(C.c, d), "device B");
        std::vector<xm_job> jobs;  // JobManager.cpp:5-13: every A segment with every B segment
        for (uint32_t i = 0; i < nsegA; i++)
            for (uint32_t j = 0; j < segs.size(); j++) jobs.push_back({i, j, sr_deg});
        xm_result* res = nullptr;
        xm_match_stats st;
        C.check(xm_match_batch(C.c, A, B, jobs.data(), jobs.size(),
                               o.fp32 ? XM_FP32_FILTER_FP64_RECHECK : XM_FP64_PLAIN,
                               o.maxout * jobs.size(), &res, &st),
                "join");
        uint64_t m = 0;
        xm_result_count(res, &m);
        total += m;
        if (!o.outdir.empty() && !write_round(C, res, m, path, o.format)) return 4;
        if (o.verbosity)
            std::cerr << "round " << r << ": " << n << " B records, " << jobs.size() << " jobs, " << m
                      << " matches, join " << st.join_ms << " ms" << std::endl;
        xm_result_destroy(res);
        xm_catalog_destroy(B);
        if (reader.joinable()) reader.join();
        read_ok = next_ok;
    }
    if (!read_ok) {
        std::cerr << "xmatch: read error on " << o.inputs[1] << std::endl;
        return 3;
    }
    for (auto& p : hb) xm_host_free(C.c, p);
    xm_catalog_destroy(A);
    close(fa);
    close(fb);
    const double secs =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
    std::cout << total << " matches in " << secs << " s";
    if (skipped) std::cout << " (" << skipped << " round(s) done before)";
    std::cout << std::endl;
    xm_ctx_destroy(C.c);
    return 0;
}
