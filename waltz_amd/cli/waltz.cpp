// waltz: command-line compressor / decompressor for raw float32 / float64
// fields on one MI355X, through the C++ API of include/waltz.hpp (the flags and
// behaviour of the reference's examples/waltz.cu:11-119).
//
//   waltz [-f|-d] -i input -z compressed -3 nx ny nz -M REL|ABS bound [-o output]
//   waltz [-f|-d] -z compressed -o output -3 nx ny nz
//
// REL bounds are resolved on the device before compression with
// WALTZ::d_extrema: abs = rel * (max - min) in double (examples/waltz.cu:50-55).
// With -o during compression the reconstruction is written and compared with
// the input (max error, PSNR, NRMSE). WALTZ_WARMUP=1 runs each API call once
// before the measured one; WALTZ_REPORT_CE2E=1 prints the API wall time and
// its throughput in the reference's convention (2^30 bytes per "GB").
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "waltz.hpp"

namespace {

void usage() {
    std::printf(
        "Usage:\n"
        "  Compress:   waltz [-f|-d] -i input -z compressed -3 nx ny nz -M REL|ABS bound [-o output]\n"
        "  Decompress: waltz [-f|-d] -z compressed -o output -3 nx ny nz\n"
        "              (-i compressed -o output also works without -M)\n"
        "\nOptions:\n"
        "  -h                  Print this help and exit.\n"
        "  -f / -d             float32 (default) / float64; -d selects the type, not decompression.\n"
        "  -i input            Raw input when compressing; compressed input when decoding without -M.\n"
        "  -z compressed       Compressed output when compressing; compressed input when decoding.\n"
        "  -o output           Write reconstruction; with compression, also verify against the input.\n"
        "  -3 nx ny nz         Original dimensions (x fastest); required for both directions.\n"
        "  -M REL|ABS bound    Compress with REL * (max - min), or an absolute error bound.\n"
        "\nTiming:\n"
        "  WALTZ_WARMUP=1: run each requested API once before the measured call.\n"
        "  WALTZ_REPORT_CE2E=1: print API wall time (excluding REL scan, file I/O and transfers).\n");
}

void check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

bool env_on(const char* name) {
    const char* v = std::getenv(name);
    return v && std::atoi(v) != 0;
}

std::vector<char> read_bytes(const char* path) {
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    if (!f) throw std::invalid_argument(std::string("Couldn't find the file: ") + path);
    const std::streamoff size = f.tellg();
    std::vector<char> out(size_t(size > 0 ? size : 0));
    f.seekg(0);
    if (size > 0 && !f.read(out.data(), size)) throw std::runtime_error(std::string("Cannot read ") + path);
    return out;
}

void write_bytes(const char* path, const void* data, size_t bytes) {
    std::ofstream f(path, std::ios::binary);
    if (!f.write(static_cast<const char*>(data), std::streamsize(bytes)))
        throw std::runtime_error(std::string("Cannot write output: ") + path);
}

// Device buffer freed on scope exit.
struct DeviceBuffer {
    void* p = nullptr;
    explicit DeviceBuffer(size_t bytes) { check(hipMalloc(&p, bytes ? bytes : 1), "hipMalloc"); }
    ~DeviceBuffer() { (void)hipFree(p); }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
};

// Wall time of one API call (the stream is synchronised by the call itself).
struct ApiClock {
    const bool on = env_on("WALTZ_REPORT_CE2E");
    const std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    void report(const char* what, size_t bytes, hipStream_t s) const {
        if (!on) return;
        (void)hipStreamSynchronize(s);
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        const double gib_s = ms > 0 ? double(bytes) / double(1ull << 30) / ms * 1000.0 : 0.0;
        std::printf("  [WALTZ-CE2E] %s_ms=%.6f throughput_GBs=%.6f\n", what, ms, gib_s);
    }
};

template <class T>
void compare(const std::vector<char>& original, const std::vector<T>& decoded) {
    const T* a = reinterpret_cast<const T*>(original.data());
    const size_t n = decoded.size();
    double lo = double(a[0]), hi = double(a[0]), max_err = 0, l2 = 0;
    size_t at = 0;
    for (size_t i = 0; i < n; ++i) {
        lo = std::min(lo, double(a[i]));
        hi = std::max(hi, double(a[i]));
        const double e = std::fabs(double(decoded[i]) - double(a[i]));
        if (e > max_err) {
            max_err = e;
            at = i;
        }
        l2 += e * e;
    }
    const double mse = l2 / double(n), range = hi - lo;
    std::printf("  Max_E  = %.15lf, idx = %zu\n", max_err, at);
    std::printf("  Max_RE = %G\n", max_err / range);
    std::printf("  PSNR = %.7lf, NRMSE = %.10G\n", 20 * std::log10(range) - 10 * std::log10(mse),
                std::sqrt(mse) / range);
}

template <class T>
void run_compress(const char* in_path, const char* cmp_path, WALTZ::Config conf) {
    const std::vector<char> host = read_bytes(in_path);
    if (host.size() != conf.num * sizeof(T)) throw std::invalid_argument("File size is not equal to the input setting");
    hipStream_t s;
    check(hipStreamCreate(&s), "hipStreamCreate");
    DeviceBuffer field(host.size());
    const size_t capacity = WALTZ::compressed_buffer_capacity<T>(conf.num);
    DeviceBuffer archive(capacity);
    check(hipMemcpy(field.p, host.data(), host.size(), hipMemcpyHostToDevice), "upload");
    if (conf.errorBoundMode == WALTZ::EB_REL) {  // outside the timed call
        double lo = 0, hi = 0;
        WALTZ::d_extrema<T>(static_cast<T*>(field.p), conf.num, lo, hi, s);
        conf.absErrorBound = conf.relErrorBound * (hi - lo);
        conf.errorBoundMode = WALTZ::EB_ABS;
        std::printf("  REL %.6g of range [%.17g, %.17g] -> ABS %.17g\n", conf.relErrorBound, lo, hi,
                    conf.absErrorBound);
    }
    if (env_on("WALTZ_WARMUP")) WALTZ::d_compress<T>(conf, static_cast<T*>(field.p), static_cast<char*>(archive.p), cmp_path, s);
    check(hipStreamSynchronize(s), "sync");
    const ApiClock clock;
    const size_t bytes = WALTZ::d_compress<T>(conf, static_cast<T*>(field.p), static_cast<char*>(archive.p), cmp_path, s);
    clock.report("compress", conf.num * sizeof(T), s);
    std::vector<char> out(bytes);
    check(hipMemcpy(out.data(), archive.p, bytes, hipMemcpyDeviceToHost), "download");
    write_bytes(cmp_path, out.data(), bytes);
    std::printf("  compressed %zu -> %zu bytes (ratio %.4f)\n", size_t(conf.num * sizeof(T)), bytes,
                double(conf.num * sizeof(T)) / double(bytes));
    (void)hipStreamDestroy(s);
}

template <class T>
void run_decompress(const char* cmp_path, const char* dec_path, const WALTZ::Config& conf, const char* original) {
    const std::vector<char> archive_host = read_bytes(cmp_path);
    if (archive_host.size() < 32) throw std::runtime_error("Cannot read compressed file or header is incomplete");
    hipStream_t s;
    check(hipStreamCreate(&s), "hipStreamCreate");
    DeviceBuffer archive(archive_host.size());
    DeviceBuffer out(conf.num * sizeof(T));
    check(hipMemcpy(archive.p, archive_host.data(), archive_host.size(), hipMemcpyHostToDevice), "upload");
    if (env_on("WALTZ_WARMUP")) WALTZ::d_decompress<T>(conf, static_cast<char*>(archive.p), static_cast<T*>(out.p), nullptr, s);
    check(hipStreamSynchronize(s), "sync");
    const ApiClock clock;
    const size_t bytes = WALTZ::d_decompress<T>(conf, static_cast<char*>(archive.p), static_cast<T*>(out.p), nullptr, s);
    clock.report("decompress", conf.num * sizeof(T), s);
    if (bytes != conf.num * sizeof(T)) throw std::runtime_error("Unexpected decompressed size");
    std::vector<T> decoded(conf.num);
    check(hipMemcpy(decoded.data(), out.p, bytes, hipMemcpyDeviceToHost), "download");
    write_bytes(dec_path, decoded.data(), bytes);
    (void)hipStreamDestroy(s);
    if (original) compare<T>(read_bytes(original), decoded);
}

// Next argument as a path unless it looks like a flag (the reference's rule
// for -z and -o, whose value may be omitted).
const char* optional_path(int argc, char** argv, int& i) {
    if (i + 1 < argc && argv[i + 1][0] != '-') return argv[++i];
    return nullptr;
}

}  // namespace

int main(int argc, char** argv) try {
    if (argc == 1) {
        usage();
        return EXIT_FAILURE;
    }
    bool want_compress = false, want_decompress = false, is_double = false;
    const char *in_path = nullptr, *cmp_path = nullptr, *dec_path = nullptr, *mode = nullptr, *bound = nullptr;
    std::vector<size_t> dims;
    for (int i = 1; i < argc; ++i) {
        const char* a = argv[i];
        if (std::strlen(a) != 2 || a[0] != '-') {
            usage();
            return EXIT_FAILURE;
        }
        const char f = a[1];
        if (f == 'h') {
            usage();
            return EXIT_SUCCESS;
        } else if (f == 'z') {
            want_compress = true;
            cmp_path = optional_path(argc, argv, i);
        } else if (f == 'o') {
            want_decompress = true;
            dec_path = optional_path(argc, argv, i);
        } else if (f == 'f' || f == 'd') {
            is_double = f == 'd';
        } else if (f == 'i') {
            if (++i == argc) return usage(), EXIT_FAILURE;
            in_path = argv[i];
        } else if (f >= '1' && f <= '4') {
            dims.clear();
            for (int k = 0; k < f - '0'; ++k) {
                size_t v = 0;
                if (++i == argc || std::sscanf(argv[i], "%zu", &v) != 1) return usage(), EXIT_FAILURE;
                dims.push_back(v);
            }
        } else if (f == 'M') {
            if (++i == argc) return usage(), EXIT_FAILURE;
            mode = argv[i];
            if (i + 1 < argc && argv[i + 1][0] != '-') bound = argv[++i];
        } else if (f == 'I') {
            std::printf("Error: data type not supported \n");
            usage();
            return EXIT_FAILURE;
        } else {
            usage();
            return EXIT_FAILURE;
        }
    }
    if (!in_path && !cmp_path) {
        std::printf("Error: you need to specify either a raw binary data file or a compressed data file as input\n");
        usage();
        return EXIT_FAILURE;
    }
    std::string tmp_archive;
    if (in_path && !cmp_path && dec_path && mode) {  // compress, decode and verify through a temporary archive
        want_compress = want_decompress = true;
        tmp_archive = std::string(in_path) + ".waltz.tmp";
        cmp_path = tmp_archive.c_str();
    }
    if (!in_path || !mode) want_compress = false;
    if ((!want_compress && !want_decompress) || (want_decompress && !dec_path) || (want_compress && !cmp_path)) {
        usage();
        return EXIT_FAILURE;
    }
    if (!want_compress && !cmp_path) cmp_path = in_path;

    WALTZ::Config conf;
    conf.setDims(dims.begin(), dims.end());
    if (conf.ndims != 3) throw std::invalid_argument("Specify the original 3D dimensions with -3 nx ny nz");
    if (mode) {
        const std::string m(mode);
        if (m == "ABS") {
            conf.errorBoundMode = WALTZ::EB_ABS;
            if (bound) conf.absErrorBound = std::atof(bound);
        } else if (m == "REL") {
            conf.errorBoundMode = WALTZ::EB_REL;
            if (bound) conf.relErrorBound = std::atof(bound);
        } else {
            std::printf("Error: wrong error bound mode setting by using the option '-M'\n");
            usage();
            return EXIT_FAILURE;
        }
    }
    if (want_compress) {
        if (is_double) run_compress<double>(in_path, cmp_path, conf);
        else run_compress<float>(in_path, cmp_path, conf);
    }
    if (want_decompress) {
        const char* original = want_compress ? in_path : nullptr;
        if (is_double) run_decompress<double>(cmp_path, dec_path, conf, original);
        else run_decompress<float>(cmp_path, dec_path, conf, original);
    }
    return EXIT_SUCCESS;
} catch (const std::exception& e) {
    std::fprintf(stderr, "waltz: %s\n", e.what());
    return EXIT_FAILURE;
}
