// dropin_driver.cpp — a C++ caller of the reference's GPU-propagation interface
// (src/celeris/cuda/propagate.hpp:16-38), compiled against include/celeris/cuda/propagate.hpp
// and linked against libceleris_rs.so exactly as the reference's analysis code would be.
// Test infrastructure: tests/test_gpu_dropin_cpp.py writes the inputs and checks the
// outputs.
//
//   dropin_driver info                    -> prints available() and device_name()
//   dropin_driver psf <in> <out>          one propagate_psf call
//   dropin_driver zx <in> <out>           one propagate_zx call
//   dropin_driver threads <in> <out>      concurrent propagate_psf calls, one per job,
//                                          each repeated `reps` times on its own thread
//
// Binary formats (little endian):
//   psf job : int32 npil, int32 n; float64 cx, cy, z, k, half_window; float64 px[npil],
//             py[npil], t[2 npil] (re, im)
//   zx job  : int32 npil, int32 nx, int32 nz; float64 xmin, xmax, zmin, zmax, k; px, py, t
//   threads : int32 njobs, int32 reps, then njobs psf jobs
//   output  : per call int32 ok, then the float64 |E|^2 grid (row-major)
#include <celeris/cuda/propagate.hpp>

#include <algorithm>
#include <atomic>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iterator>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace {

struct Reader {
    std::vector<char> buf;
    size_t pos = 0;
    template <class T>
    T get() {
        T v;
        if (pos + sizeof(T) > buf.size()) throw std::runtime_error("input too short");
        std::memcpy(&v, buf.data() + pos, sizeof(T));
        pos += sizeof(T);
        return v;
    }
    void get_doubles(std::vector<double>& v, size_t n) {
        v.resize(n);
        for (size_t i = 0; i < n; ++i) v[i] = get<double>();
    }
};

struct PsfJob {
    int npil = 0, n = 0;
    double cx = 0, cy = 0, z = 0, k = 0, W = 0;
    std::vector<double> px, py, t;
    std::vector<double> out;
    int ok = 0;
    void read(Reader& r) {
        npil = r.get<int32_t>();
        n = r.get<int32_t>();
        cx = r.get<double>();
        cy = r.get<double>();
        z = r.get<double>();
        k = r.get<double>();
        W = r.get<double>();
        r.get_doubles(px, npil);
        r.get_doubles(py, npil);
        r.get_doubles(t, 2 * static_cast<size_t>(npil));
    }
    void run() {
        out.assign(static_cast<size_t>(std::max(n, 0)) * std::max(n, 0), -1.0);  // -1: untouched
        const auto* pt = reinterpret_cast<const std::complex<double>*>(t.data());
        ok = celeris::cuda::propagate_psf(px.data(), py.data(), pt, npil, cx, cy, z, k, n, W,
                                          out.data())
                 ? 1
                 : 0;
    }
};

Reader read_file(const char* path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error(std::string("cannot open ") + path);
    Reader r;
    r.buf.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
    return r;
}

void write_results(const char* path, const std::vector<std::pair<int, const std::vector<double>*>>& res) {
    std::ofstream f(path, std::ios::binary);
    for (const auto& [ok, v] : res) {
        const int32_t k = ok;
        f.write(reinterpret_cast<const char*>(&k), sizeof k);
        f.write(reinterpret_cast<const char*>(v->data()), static_cast<std::streamsize>(v->size() * sizeof(double)));
    }
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: dropin_driver info | psf|zx|threads <in> <out>\n");
        return 2;
    }
    const std::string cmd = argv[1];
    try {
        if (cmd == "info") {
            std::printf("available %d\ndevice %s\n", celeris::cuda::available() ? 1 : 0,
                        celeris::cuda::device_name());
            return 0;
        }
        if (argc < 4) return 2;
        Reader r = read_file(argv[2]);
        if (cmd == "psf") {
            PsfJob j;
            j.read(r);
            j.run();
            write_results(argv[3], {{j.ok, &j.out}});
        } else if (cmd == "zx") {
            const int npil = r.get<int32_t>(), nx = r.get<int32_t>(), nz = r.get<int32_t>();
            const double xmin = r.get<double>(), xmax = r.get<double>(), zmin = r.get<double>(),
                         zmax = r.get<double>(), k = r.get<double>();
            std::vector<double> px, py, t;
            r.get_doubles(px, npil);
            r.get_doubles(py, npil);
            r.get_doubles(t, 2 * static_cast<size_t>(npil));
            std::vector<double> out(static_cast<size_t>(std::max(nx, 0)) * std::max(nz, 0), -1.0);
            const int ok = celeris::cuda::propagate_zx(
                               px.data(), py.data(), reinterpret_cast<const std::complex<double>*>(t.data()),
                               npil, xmin, xmax, nx, zmin, zmax, nz, k, out.data())
                               ? 1
                               : 0;
            write_results(argv[3], {{ok, &out}});
        } else if (cmd == "threads") {
            const int njobs = r.get<int32_t>(), reps = r.get<int32_t>();
            std::vector<PsfJob> jobs(njobs);
            for (PsfJob& j : jobs) j.read(r);
            std::atomic<int> ready{0};
            std::vector<std::thread> th;
            std::vector<int> all_ok(njobs, 1);
            for (int i = 0; i < njobs; ++i)
                th.emplace_back([&, i] {
                    ready.fetch_add(1);
                    while (ready.load() < njobs) std::this_thread::yield();
                    for (int rep = 0; rep < reps; ++rep) {
                        jobs[i].run();
                        all_ok[i] &= jobs[i].ok;
                    }
                });
            for (std::thread& t : th) t.join();
            std::vector<std::pair<int, const std::vector<double>*>> res;
            for (int i = 0; i < njobs; ++i) res.push_back({all_ok[i], &jobs[i].out});
            write_results(argv[3], res);
        } else {
            return 2;
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "dropin_driver: %s\n", e.what());
        return 1;
    }
    return 0;
}
