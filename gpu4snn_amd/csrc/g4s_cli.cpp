// g4s_cli.cpp -- the `gpu4snn N Nsyn iterations` executable.
//
// Reference-compatible front end of main() in gpu4snn.cu (cu:735-1332): by
// default it simulates every (regime, algorithm) pair -- 3 regimes x 5
// schemes (cu:1082-1083) -- on one network and writes, per run,
// Results/N{N}Nsyn{Nsyn}Regime{mode}Alg{alg}/ with the reference's files:
//   neuroninfo.csv       a b c d v u I id nospks   (initial state, cu:1120-1127)
//   spiketimes.csv       SpkTime NeuronID         (cu:1134, cu:1256-1261)
//   computetime.csv      TimeStep SpksPerStep TimeKernelUpdate TimeSpent (cu:1140, cu:1267)
//   sim_overview.csv     N elapsed(ms) elapsedwithdata(ms) totalspkcount, no
//                        trailing newline (cu:1311-1313)
//   neuroninfoafter.csv  the final state (cu:1318-1327; the reference wrote
//                        the pre-simulation copy there)
// Optional flags select a subset or change the accumulation mode/seed.
// Spike rows are streamed to spiketimes.csv chunk by chunk while the device
// runs the next chunk (g4s_run_stream).  Per-step times in computetime.csv
// are device times of each step (g4s_get_step_times), and after each run
// the dominant kernel's launch shape and occupancy is printed, as the
// reference prints its kernels' (cu:800-838).
#include "gpu4snn.h"

#include <sys/stat.h>

#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

namespace {

const char *kAlgNames[] = {"AP", "N", "S", "AB", "SKL"};

void usage()
{
    std::cerr << "usage: gpu4snn N Nsyn iterations [--alg AP|N|S|AB|SKL|all] [--regime 0|1|2|all]\n"
                 "                [--mode atomic|det] [--seed S] [--scale X] [--outdir DIR]\n";
}

int parse_alg(const char *s)
{
    for (int k = 0; k < 5; ++k)
        if (!strcmp(s, kAlgNames[k])) return k;
    char *end = nullptr;
    long v = strtol(s, &end, 10);
    if (end && *end == 0 && v >= 0 && v <= 4) return (int)v;
    return -2;
}

void write_neuroninfo(const std::string &path, int64_t n, const std::vector<float> &a, const std::vector<float> &b,
                      const std::vector<float> &c, const std::vector<float> &d, const std::vector<float> &v,
                      const std::vector<float> &u, const std::vector<float> &I, const std::vector<int32_t> &nos,
                      bool *ok)
{
    std::ofstream f(path.c_str());
    if (!f.is_open()) {
        *ok = false;
        return;
    }
    f << "a b c d v u I id nospks" << "\n";
    for (int64_t i = 0; i < n; ++i)
        f << a[i] << " " << b[i] << " " << c[i] << " " << d[i] << " " << v[i] << " " << u[i] << " " << I[i] << " "
          << i << " " << nos[i] << "\n";
    *ok = true;
}

// Streaming sink: spike rows of each chunk go to spiketimes.csv while the
// device runs the next chunk; counts are kept for computetime.csv.
struct Sink {
    std::ofstream *spikes;
    std::vector<int32_t> counts;
    int64_t words;
    long long total;
};

int write_chunk(void *user, int64_t t0, int32_t rows, const uint32_t *raster, const int32_t *counts)
{
    Sink *k = (Sink *)user;
    std::string buf;
    buf.reserve(4096);
    char line[48];
    for (int32_t r = 0; r < rows; ++r) {
        const uint32_t *row = raster + (size_t)r * (size_t)k->words;
        for (int64_t w = 0; w < k->words; ++w) {
            uint32_t bits = row[w];
            while (bits) {
                const int b = __builtin_ctz(bits);
                const int len = snprintf(line, sizeof(line), "%lld %lld\n", (long long)(t0 + r), (long long)(w * 32 + b));
                buf.append(line, (size_t)len);
                bits &= bits - 1;
            }
        }
        k->counts.push_back(counts[r]);
        k->total += counts[r];
    }
    k->spikes->write(buf.data(), (std::streamsize)buf.size());
    return k->spikes->good() ? 0 : 1;
}

int check(int rc, const char *what)
{
    if (rc != G4S_OK) {
        std::cerr << "gpu4snn: " << what << " failed (" << rc << "): " << g4s_last_error() << "\n";
        exit(2);
    }
    return rc;
}

}  // namespace

int main(int argc, char **argv)
{
    if (argc < 4) {
        usage();
        return 2;
    }
    const long long N = atoll(argv[1]);
    const long nsyn = atol(argv[2]);
    const long iters = atol(argv[3]);
    int alg_sel = -1, regime_sel = -1, mode = G4S_ACCUM_ATOMIC;
    unsigned long long seed = 20221;
    float scale = 1.0f;
    std::string outdir = "Results";
    for (int k = 4; k < argc; ++k) {
        std::string a = argv[k];
        auto next = [&]() -> const char * {
            if (k + 1 >= argc) { usage(); exit(2); }
            return argv[++k];
        };
        if (a == "--alg") {
            const char *v = next();
            alg_sel = strcmp(v, "all") ? parse_alg(v) : -1;
            if (alg_sel == -2) { usage(); return 2; }
        } else if (a == "--regime") {
            const char *v = next();
            regime_sel = strcmp(v, "all") ? atoi(v) : -1;
        } else if (a == "--mode") {
            const char *v = next();
            if (!strcmp(v, "det")) mode = G4S_ACCUM_DET;
            else if (!strcmp(v, "atomic")) mode = G4S_ACCUM_ATOMIC;
            else { usage(); return 2; }
        } else if (a == "--seed") {
            seed = strtoull(next(), nullptr, 10);
        } else if (a == "--scale") {
            scale = (float)atof(next());
        } else if (a == "--outdir") {
            outdir = next();
        } else {
            usage();
            return 2;
        }
    }
    if (N < 2 || nsyn < 1 || nsyn >= N || iters < 0) {
        std::cerr << "gpu4snn: need N >= 2, 1 <= Nsyn < N, iterations >= 0\n";
        return 2;
    }
    const int64_t n_exc = (4 * N) / 5;
    std::cout << "Number of neurons: " << N << "\nnumber of synapses per neuron: " << nsyn << "\n";
    printf("%lld Neurons in the network: %lld excitatory and %lld inhibitory \n", N, (long long)n_exc,
           (long long)(N - n_exc));

    g4s_spec spec{};
    spec.n = N;
    spec.nsyn = (int32_t)nsyn;
    spec.regime = regime_sel >= 0 ? regime_sel : 0;
    spec.seed = seed;
    spec.thal_scale = scale;
    spec.n_exc = -1;
    g4s_net *net = nullptr;
    printf("Building the network and uploading it to the GPU\n");
    check(g4s_create(&spec, &net), "g4s_create");

    const int64_t words = (N + 31) / 32;
    g4s_run_opts ropts{};
    ropts.chunk = 1000;
    std::vector<float> a(N), b(N), c(N), d(N), v(N), u(N), I(N);
    std::vector<int32_t> nos(N);
    mkdir(outdir.c_str(), S_IRWXU | S_IRWXG | S_IROTH | S_IXOTH);

    for (int regime = 0; regime < 3; ++regime) {
        if (regime_sel >= 0 && regime != regime_sel) continue;
        for (int alg = 0; alg < 5; ++alg) {
            if (alg_sel >= 0 && alg != alg_sel) continue;
            check(g4s_set_regime(net, regime, scale), "g4s_set_regime");
            check(g4s_reset(net), "g4s_reset");
            std::cout << "===================================================== \n";
            std::cout << "Neurons: " << N << " Synapses: " << nsyn << " Dynamic (Algorithm): " << alg
                      << " Mode (State): " << regime << "\n";
            std::cout << "===================================================== \n";
            char buf[256];
            snprintf(buf, sizeof(buf), "%s/N%lldNsyn%ldRegime%dAlg%d/", outdir.c_str(), N, nsyn, regime, alg);
            const std::string dir = buf;
            mkdir(dir.c_str(), S_IRWXU | S_IRWXG | S_IROTH | S_IXOTH);

            check(g4s_get_params(net, a.data(), b.data(), c.data(), d.data()), "g4s_get_params");
            check(g4s_get_state(net, v.data(), u.data(), I.data(), nos.data()), "g4s_get_state");
            bool ok = false;
            write_neuroninfo(dir + "neuroninfo.csv", N, a, b, c, d, v, u, I, nos, &ok);
            std::ofstream spikes((dir + "spiketimes.csv").c_str());
            spikes << "SpkTime NeuronID" << "\n";
            std::ofstream ctime((dir + "computetime.csv").c_str());
            ctime << "TimeStep SpksPerStep TimeKernelUpdate TimeSpent" << "\n";

            std::cout << "\nStarting simulation and timer...\n";
            Sink sink{&spikes, {}, words, 0};
            sink.counts.reserve((size_t)iters);
            if (iters > 0)
                check(g4s_run_stream(net, alg, mode, (int32_t)iters, write_chunk, &sink, &ropts), "g4s_run_stream");
            g4s_timing tm{};
            g4s_get_timing(net, &tm);
            const double dev_ms = iters > 0 ? tm.device_ms : 0.0, wall_ms = iters > 0 ? tm.wall_ms : 0.0;
            const long long total = sink.total;
            // TimeKernelUpdate: the step's device time (start to start of
            // consecutive steps on the device clock); TimeSpent adds the
            // step's share of what the run spent outside the device (chunk
            // copies, spiketimes.csv writes).
            std::vector<float> step_ms((size_t)std::max<long>(iters, 0L), 0.0f);
            if (iters > 0 && g4s_get_step_times(net, step_ms.data(), (int32_t)iters) < 0)
                check(G4S_ERR_STATE, "g4s_get_step_times");
            const double host_per = iters > 0 ? std::max(0.0, wall_ms - dev_ms) / (double)iters : 0.0;
            for (size_t r = 0; r < sink.counts.size(); ++r)
                ctime << r << " " << sink.counts[r] << " " << step_ms[r] << " " << step_ms[r] + host_per << "\n";
            if (iters > 0) {
                g4s_launch_info li{};
                g4s_get_launch_info(net, &li);
                printf("Launched %s: grid %d, block %d, LDS %d bytes per workgroup, %d VGPRs, %d SGPRs, "
                       "%d waves per SIMD, %d workgroups per CU of %d CUs%s\n",
                       li.kernel, li.grid, li.block, li.lds_bytes, li.vgprs, li.sgprs, li.waves_per_simd,
                       li.resident_per_cu, li.cus, li.persistent ? " (persistent: grid within residency)" : "");
            }
            std::cout << "\nEnd of simulation...\n";
            std::cout << "Simulation took: " << wall_ms << " ms.\n";
            std::cout << "Without data transfers: " << dev_ms << " ms.\n";
            std::cout << "Total number of spikes: " << total << "\n\n\n";
            std::ofstream ov((dir + "sim_overview.csv").c_str());
            ov << "N elapsed(ms) elapsedwithdata(ms) totalspkcount" << "\n";
            ov << N << " " << dev_ms << " " << wall_ms << " " << total;
            ov.close();

            check(g4s_get_state(net, v.data(), u.data(), I.data(), nos.data()), "g4s_get_state");
            write_neuroninfo(dir + "neuroninfoafter.csv", N, a, b, c, d, v, u, I, nos, &ok);
            if (!ok) {
                std::cerr << "Error! Failed to open file " << std::endl;
                g4s_destroy(net);
                return 1;
            }
        }
    }
    g4s_destroy(net);
    return 0;
}
