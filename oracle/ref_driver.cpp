// ref_driver.cpp -- command-line driver around the reference's own functions.
//
// TEST INFRASTRUCTURE ONLY.  oracle/build_ref.sh links this file against the
// reference sources where they lie (src/common/{spgemm_common,matrix_market,
// full_reference}.cpp, src/baselines/spgemm_exchange.cpp,
// src/trident/common/matrix_blocks.cpp) into oracle/_ref/refdrv.  It is used to
// pin the C restatement in oracle.c, to mint the golden fixtures under
// tests/golden/, and as the CPU baseline ("kind": "reference") of bench.py.
//
// Functions called (reference file:line):
//   readMatrixMarket            src/common/matrix_market.cpp:28-144
//   spgemm (-np 1, empty halo)  src/baselines/spgemm_exchange.cpp:258-325
//   serialSpgemm                src/common/spgemm_common.cpp:470-472
//   buildPartialReference       src/common/spgemm_common.cpp:541-554
//   scalarMultiplicationCount   src/common/spgemm_common.cpp:474-487
//   percentile90                src/common/spgemm_common.cpp:916-924
//   trident::LocalAccumulator   src/trident/common/matrix_blocks.cpp:152-200
// The timing loop restates src/baselines/two_sided_spgemm.cpp:126-167 for one
// rank (omp_set_schedule as at :70-72), for one product (bench) or for a
// sequence of independent products timed as one sample (benchjobs).
//
// Binary CSR files (*.csr): "SPGCSR01", int32 rows, int32 cols, int64 nnz,
// int64 row_ptr[rows+1], int32 col[nnz], float64 val[nnz], little endian.
#include "matrix_market.hpp"
#include "spgemm_exchange.hpp"
#include "matrix_blocks.hpp"

#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

bool endsWith(const std::string& s, const std::string& suffix) {
    return s.size() >= suffix.size() && s.compare(s.size() - suffix.size(), suffix.size(), suffix) == 0;
}

CsrMatrix readCsrFile(const std::string& path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error("cannot open " + path);
    char magic[8];
    in.read(magic, 8);
    if (std::memcmp(magic, "SPGCSR01", 8) != 0) throw std::runtime_error("bad csr magic: " + path);
    CsrMatrix m;
    std::int64_t nnz = 0;
    in.read(reinterpret_cast<char*>(&m.rows), 4);
    in.read(reinterpret_cast<char*>(&m.cols), 4);
    in.read(reinterpret_cast<char*>(&nnz), 8);
    m.rowPtr.resize(static_cast<std::size_t>(m.rows) + 1);
    m.columnIndices.resize(static_cast<std::size_t>(nnz));
    m.values.resize(static_cast<std::size_t>(nnz));
    in.read(reinterpret_cast<char*>(m.rowPtr.data()), static_cast<std::streamsize>(m.rowPtr.size() * 8));
    in.read(reinterpret_cast<char*>(m.columnIndices.data()), static_cast<std::streamsize>(nnz * 4));
    in.read(reinterpret_cast<char*>(m.values.data()), static_cast<std::streamsize>(nnz * 8));
    if (!in) throw std::runtime_error("truncated csr file: " + path);
    return m;
}

void writeCsrFile(const std::string& path, const CsrMatrix& m) {
    std::ofstream out(path, std::ios::binary);
    const std::int64_t nnz = static_cast<std::int64_t>(m.values.size());
    out.write("SPGCSR01", 8);
    out.write(reinterpret_cast<const char*>(&m.rows), 4);
    out.write(reinterpret_cast<const char*>(&m.cols), 4);
    out.write(reinterpret_cast<const char*>(&nnz), 8);
    out.write(reinterpret_cast<const char*>(m.rowPtr.data()), static_cast<std::streamsize>(m.rowPtr.size() * 8));
    out.write(reinterpret_cast<const char*>(m.columnIndices.data()), static_cast<std::streamsize>(nnz * 4));
    out.write(reinterpret_cast<const char*>(m.values.data()), static_cast<std::streamsize>(nnz * 8));
    if (!out) throw std::runtime_error("cannot write " + path);
}

CsrMatrix load(const std::string& path) {
    return endsWith(path, ".csr") ? readCsrFile(path) : readMatrixMarket(path);
}

CsrMatrix localSpgemm(const CsrMatrix& a, const CsrMatrix& b) {
    RemoteSparseRows halo;  // -np 1: every B row is local
    return spgemm(a, RowBlock{0, b.rows}, b, halo, MPI_COMM_WORLD);
}

CsrMatrix firstRows(const CsrMatrix& a, int rows) {
    if (rows >= a.rows) return a;
    CsrMatrix s;
    s.rows = rows;
    s.cols = a.cols;
    s.rowPtr.assign(a.rowPtr.begin(), a.rowPtr.begin() + rows + 1);
    s.columnIndices.assign(a.columnIndices.begin(), a.columnIndices.begin() + s.rowPtr.back());
    s.values.assign(a.values.begin(), a.values.begin() + s.rowPtr.back());
    return s;
}

int usage() {
    std::cerr << "usage: refdrv read A out.csr | spgemm A B out.csr | serial A B out.csr |\n"
                 "       partial A B out.bin | count A B |\n"
                 "       accumulate out.csr rows cols A1 B1 [A2 B2 ...] |\n"
                 "       bench A B threads warmup repeats trials [row_limit] |\n"
                 "       benchjobs manifest threads warmup repeats trials\n"
                 "(manifest lines: \"ab A B\" for A*B, \"rap R A P\" for (R*A)*P)\n";
    return 2;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) return usage();
    const std::string cmd = argv[1];
    omp_set_dynamic(0);
    omp_set_schedule(omp_sched_guided, 64);
    try {
        if (cmd == "read" && argc == 4) {
            writeCsrFile(argv[3], readMatrixMarket(argv[2]));
        } else if (cmd == "spgemm" && argc == 5) {
            writeCsrFile(argv[4], localSpgemm(load(argv[2]), load(argv[3])));
        } else if (cmd == "serial" && argc == 5) {
            writeCsrFile(argv[4], serialSpgemm(load(argv[2]), load(argv[3])));
        } else if (cmd == "partial" && argc == 5) {
            const PartialReference ref = buildPartialReference(load(argv[2]), load(argv[3]));
            std::ofstream out(argv[4], std::ios::binary);
            const std::int32_t finite = ref.finite ? 1 : 0;
            out.write(reinterpret_cast<const char*>(&finite), 4);
            out.write(reinterpret_cast<const char*>(ref.rowNnz.data()),
                      static_cast<std::streamsize>(ref.rowNnz.size() * 8));
        } else if (cmd == "count" && argc == 4) {
            std::cout << scalarMultiplicationCount(load(argv[2]), load(argv[3])) << '\n';
        } else if (cmd == "accumulate" && argc >= 7 && (argc - 5) % 2 == 0) {
            trident::LocalAccumulator acc(std::stoi(argv[3]), std::stoi(argv[4]));
            for (int i = 5; i < argc; i += 2) acc.addProduct(load(argv[i]), load(argv[i + 1]));
            writeCsrFile(argv[2], acc.finish());
        } else if (cmd == "bench" && (argc == 8 || argc == 9)) {
            const int threads = std::stoi(argv[4]);
            const int warmup = std::stoi(argv[5]);
            const int repeats = std::stoi(argv[6]);
            const int trials = std::stoi(argv[7]);
            CsrMatrix a = load(argv[2]);
            const CsrMatrix b = load(argv[3]);
            if (argc == 9) a = firstRows(a, std::stoi(argv[8]));
            omp_set_num_threads(threads);
            CsrMatrix c = localSpgemm(a, b);  // first product
            for (int i = 0; i < warmup; ++i) c = localSpgemm(a, b);
            std::vector<double> samples;
            for (int t = 0; t < trials; ++t) {
                for (int r = 0; r < repeats; ++r) {
                    const auto start = std::chrono::steady_clock::now();
                    c = localSpgemm(a, b);
                    samples.push_back(std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count());
                }
            }
            std::vector<double> sorted = samples;
            std::sort(sorted.begin(), sorted.end());
            const double p90 = percentile90(samples);
            const double p50 = sorted[(sorted.size() - 1) / 2];
            const std::int64_t products = scalarMultiplicationCount(a, b);
            std::printf("{\"rows\": %d, \"products\": %lld, \"c_nnz\": %zu, \"threads\": %d, "
                        "\"compute_p90_s\": %.9g, \"compute_p50_s\": %.9g, \"gflops_p90\": %.6g, "
                        "\"samples\": %zu}\n",
                        a.rows, static_cast<long long>(products), c.values.size(), threads, p90, p50,
                        p90 > 0 ? 2.0 * static_cast<double>(products) / p90 / 1e9 : 0.0, samples.size());
        } else if (cmd == "benchjobs" && argc == 7) {
            // A sequence of independent products timed as one sample, with
            // the timing loop of two_sided_spgemm.cpp:126-167 (first run,
            // warmup, trials x repeats, nearest-rank P90).
            const int threads = std::stoi(argv[3]);
            const int warmup = std::stoi(argv[4]);
            const int repeats = std::stoi(argv[5]);
            const int trials = std::stoi(argv[6]);
            struct Job {
                CsrMatrix a, b, p;
                bool triple;
            };
            std::vector<Job> jobs;
            std::ifstream manifest(argv[2]);
            std::string kind;
            while (manifest >> kind) {
                Job j;
                std::string pa, pb, pp;
                if (kind == "ab") {
                    manifest >> pa >> pb;
                    j.triple = false;
                } else if (kind == "rap") {
                    manifest >> pa >> pb >> pp;
                    j.triple = true;
                    j.p = load(pp);
                } else {
                    throw std::runtime_error("bad manifest line kind: " + kind);
                }
                j.a = load(pa);
                j.b = load(pb);
                jobs.push_back(std::move(j));
            }
            omp_set_num_threads(threads);
            std::int64_t products = 0, cnnz = 0;
            auto pass = [&](bool count) {
                for (const Job& j : jobs) {
                    CsrMatrix c = localSpgemm(j.a, j.b);
                    if (count) products += scalarMultiplicationCount(j.a, j.b);
                    if (j.triple) {
                        if (count) products += scalarMultiplicationCount(c, j.p);
                        c = localSpgemm(c, j.p);
                    }
                    if (count) cnnz += static_cast<std::int64_t>(c.values.size());
                }
            };
            pass(true);  // first run
            for (int i = 0; i < warmup; ++i) pass(false);
            std::vector<double> samples;
            for (int t = 0; t < trials; ++t) {
                for (int r = 0; r < repeats; ++r) {
                    const auto start = std::chrono::steady_clock::now();
                    pass(false);
                    samples.push_back(std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count());
                }
            }
            std::vector<double> sorted = samples;
            std::sort(sorted.begin(), sorted.end());
            const double p90 = percentile90(samples);
            const double p50 = sorted[(sorted.size() - 1) / 2];
            std::printf("{\"jobs\": %zu, \"products\": %lld, \"c_nnz\": %lld, \"threads\": %d, "
                        "\"p90_s\": %.9g, \"p50_s\": %.9g, \"jobs_per_s_p90\": %.6g, \"gflops_p90\": %.6g, "
                        "\"samples\": %zu}\n",
                        jobs.size(), static_cast<long long>(products), static_cast<long long>(cnnz), threads, p90,
                        p50, p90 > 0 ? static_cast<double>(jobs.size()) / p90 : 0.0,
                        p90 > 0 ? 2.0 * static_cast<double>(products) / p90 / 1e9 : 0.0, samples.size());
        } else {
            return usage();
        }
    } catch (const std::exception& e) {
        std::cerr << "refdrv: " << e.what() << '\n';
        return 1;
    }
    return 0;
}
