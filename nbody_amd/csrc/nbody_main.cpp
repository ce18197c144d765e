// nbody — command-line driver over the HIP algorithm table, with the
// reference driver's flags, messages, cycling and output line
// (src/nbody.cu:735-1122), so its runs and `scripts/nbody-stats.py` parsing
// carry over. Differences: every algorithm runs on the GPU (the CPU table
// rows are the reference's, not this library's); the crosscheck golden is the
// library's FP64 direct sum unless --golden=cpu_soa asks for the reference's
// CPU_SOA sum; positions and velocities stay on the device and are integrated
// there; there is no keyboard control or OpenGL window.
#include "nbody_hip.h"
#include "nbody_ics.h"

#include <hip/hip_runtime.h>

#include <getopt.h>
#include <strings.h>

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

namespace {

enum AlgoType { ALGO_AOS = 1, ALGO_AOS_GPU = 2 };  // src/nbody.h:58-64

typedef float (*pfnForce)(float*, float*, float, size_t);

struct AlgoDef {
    const char* name;
    AlgoType type;
    pfnForce fn;
};

// The table of src/nbody.cu:83-103, filled with this library's entries.
const AlgoDef kAlgorithms[] = {
    {"HIP_AOS", ALGO_AOS_GPU, ComputeGravitation_HIP_AOS},
    {"HIP_LDS", ALGO_AOS_GPU, ComputeGravitation_HIP_LDS},
    {"HIP_DPP", ALGO_AOS_GPU, ComputeGravitation_HIP_DPP},
    {"HIP_SGPR", ALGO_AOS_GPU, ComputeGravitation_HIP_SGPR},
    {"HIP_Symmetric", ALGO_AOS_GPU, ComputeGravitation_HIP_Symmetric},
    {"HIP", ALGO_AOS_GPU, ComputeGravitation_HIP},
    {"HIP_Host", ALGO_AOS, ComputeGravitation_HIP_Host},
    {nullptr, ALGO_AOS, nullptr},
};

const int kDefaultKParticles = 16;  // src/nbody.cu:80
const unsigned kDefaultSeed = 7;    // src/nbody.cu:81
const float kSoftening = 0.1f;      // src/nbody.cu:151-156
const float kDamping = 0.95f;
const float kTimestep = 0.016f;
const float kScale = 0.9f;
const float kVelocityScale = 8.0f;

int g_maxGPUs = 0;
int g_numGPUs = 0;
int g_bNoCPU = 0;
int g_bCrossCheck = 1;
int g_bGoldenSOA = 0;  // --golden=cpu_soa

int maxAlgorithmIdx()
{
    int idx = 0;
    while (kAlgorithms[idx].name) idx++;
    return idx - 1;
}

void print_algorithms()
{
    const bool available = g_maxGPUs > 0, enabled = g_numGPUs > 0;
    fprintf(stderr, "\nAlgorithms available in this build:\n\n");
    for (int idx = 0; kAlgorithms[idx].name; idx++) {
        const char* suffix = "";
        if (!available) suffix = " [disabled, no GPUs available]";
        else if (!enabled) suffix = " [disabled by user]";
        else if (g_bNoCPU && kAlgorithms[idx].type == ALGO_AOS) suffix = " [disabled by user]";
        fprintf(stdout, "   %d - %s%s\n", idx, kAlgorithms[idx].name, suffix);
    }
    fprintf(stderr, "\n");
}

void print_usage()
{
    fprintf(stderr, "Usage: nbody [arguments]\n\nAvailable arguments:\n\n");
    fprintf(stderr, "    --bodies=<N> | -n <N>\n        Number of bodies, multiplied by 1024. [default: %d]\n\n",
            kDefaultKParticles);
    fprintf(stderr, "    --gpus=<N> | -g <N>\n        GPUs to use (this build computes on one).\n"
                    "        [default: number of available GPUs]\n\n");
    fprintf(stderr, "    --no-cpu\n        Disables the host-pointer algorithm and the crosscheck.\n\n");
    fprintf(stderr, "    --no-crosscheck\n        Disables cross-validation against the golden forces.\n\n");
    fprintf(stderr, "    --golden=<fp64|cpu_soa>\n        Crosscheck golden: the FP64 direct sum on the GPU, or the\n"
                    "        reference's CPU_SOA sum on the host's threads. [default: fp64]\n\n");
    fprintf(stderr, "    --iterations=<N> | -i <N>\n        Passes through the algorithm list. [default: loop forever]\n\n");
    fprintf(stderr, "    --cycle-after=<N> | -c <N>\n        Steps before moving to the next algorithm.\n"
                    "        [default: none, don't cycle]\n\n");
    fprintf(stderr, "    --seed=<N> | -s <N>\n        Seed of the initial conditions. [default: %u]\n\n", kDefaultSeed);
    fprintf(stderr, "    --ic=<reference|ball|plummer>\n        Initial conditions: the reference's randomUnitBodies,\n"
                    "        a uniform ball of radius 1.54 (N/16384)^(1/3), or a Plummer sphere. [default: reference]\n\n");
    fprintf(stderr, "    --list\n        Lists the available algorithms.\n\n");
    fprintf(stderr, "    --algorithm=<name or index> | -a <name or index>\n        Algorithm to start with. [default: %s]\n\n",
            kAlgorithms[0].name);
    fprintf(stderr, "    --verbose\n        Print every step instead of at most ten lines per second.\n\n");
    fprintf(stderr, "    --help\n        Prints this help text.\n");
}

#define HIP_OK(x)                                                                                     \
    do {                                                                                              \
        hipError_t e_ = (x);                                                                          \
        if (e_ != hipSuccess) {                                                                       \
            fprintf(stderr, "HIP error: %s (%s)\n", hipGetErrorString(e_), #x);                       \
            return 1;                                                                                 \
        }                                                                                             \
    } while (0)

struct State {
    size_t n = 0;
    float *d_pm = nullptr, *d_vm = nullptr, *d_force = nullptr;
    double* d_golden = nullptr;
    std::vector<float> h_pm, h_force, golden;
    std::vector<double> h_golden;
};

// src/nbody.cu:120-125
float relError(float a, float b)
{
    if (a == b) return 0.0f;
    return fabsf(a - b) / b;
}

// The golden of src/nbody.cu:311-321 with --golden=cpu_soa: CPU_SOA's FP32
// sum (src/nbody_CPU_SOA.cpp:68-96, the pair term of
// src/bodybodyInteraction.cuh:40-65 with rsqrtf as 1/sqrtf, src/chCUDA.h:143-146)
// over host positions. Bodies i are split over the host's threads; j runs in
// 16 interleaved partial sums added at the end, the shape of a 16-lane
// `omp simd` reduction, so the result equals the reference's build up to the
// order in which its reduction adds those lanes.
void golden_cpu_soa(const float* pm, float eps2, size_t n, float* golden)
{
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const size_t nthreads = std::min<size_t>(hw, (n + 255) / 256);
    auto rows = [&](size_t t) {
        for (size_t i = t; i < n; i += nthreads) {
            const float xi = pm[4 * i + 0], yi = pm[4 * i + 1], zi = pm[4 * i + 2];
            float ax[16] = {}, ay[16] = {}, az[16] = {};
            for (size_t j0 = 0; j0 < n; j0 += 16) {
                const size_t m = std::min<size_t>(16, n - j0);
                for (size_t l = 0; l < m; l++) {
                    const float* b = pm + 4 * (j0 + l);
                    const float dx = b[0] - xi, dy = b[1] - yi, dz = b[2] - zi;
                    float d2 = dx * dx + dy * dy + dz * dz;
                    d2 += eps2;
                    const float inv = 1.0f / sqrtf(d2);
                    const float s = b[3] * (inv * inv * inv);
                    ax[l] += dx * s;
                    ay[l] += dy * s;
                    az[l] += dz * s;
                }
            }
            float sx = 0.0f, sy = 0.0f, sz = 0.0f;
            for (int l = 0; l < 16; l++) {
                sx += ax[l];
                sy += ay[l];
                sz += az[l];
            }
            golden[4 * i + 0] = sx;
            golden[4 * i + 1] = sy;
            golden[4 * i + 2] = sz;
            golden[4 * i + 3] = 0.0f;
        }
    };
    std::vector<std::thread> pool;
    for (size_t t = 1; t < nthreads; t++) pool.emplace_back(rows, t);
    rows(0);
    for (auto& th : pool) th.join();
}

// One step of src/nbody.cu:293-470: golden, force, crosscheck, integrate.
// Returns nonzero when the algorithm was skipped or failed.
int ComputeGravitation(State& s, float* ms, float* maxRelError, const AlgoDef& alg, int crossCheck)
{
    if (g_numGPUs < 1) return 1;
    if (g_bNoCPU && alg.type == ALGO_AOS) return 1;
    const size_t n = s.n;
    const float eps2 = kSoftening * kSoftening;
    if (crossCheck && g_bGoldenSOA) {
        HIP_OK(hipMemcpy(s.h_pm.data(), s.d_pm, 4 * n * sizeof(float), hipMemcpyDeviceToHost));
        golden_cpu_soa(s.h_pm.data(), eps2, n, s.golden.data());
    } else if (crossCheck) {
        if (nbody_hip_force_fp64(s.d_golden, s.d_pm, (double)eps2, n) <= 0.0f) {
            fprintf(stderr, "FP64 golden failed: %s\n", nbody_hip_last_error());
            return 1;
        }
        HIP_OK(hipMemcpy(s.h_golden.data(), s.d_golden, 4 * n * sizeof(double), hipMemcpyDeviceToHost));
        for (size_t k = 0; k < 4 * n; k++) s.golden[k] = (float)s.h_golden[k];
    }
    HIP_OK(hipMemset(s.d_force, 0, 4 * n * sizeof(float)));
    if (alg.type == ALGO_AOS_GPU) {
        *ms = alg.fn(s.d_force, s.d_pm, eps2, n);
        if (*ms >= FLT_EPSILON)
            HIP_OK(hipMemcpy(s.h_force.data(), s.d_force, 4 * n * sizeof(float), hipMemcpyDeviceToHost));
    } else {
        HIP_OK(hipMemcpy(s.h_pm.data(), s.d_pm, 4 * n * sizeof(float), hipMemcpyDeviceToHost));
        std::fill(s.h_force.begin(), s.h_force.end(), 0.0f);
        *ms = alg.fn(s.h_force.data(), s.h_pm.data(), eps2, n);
        if (*ms >= FLT_EPSILON)
            HIP_OK(hipMemcpy(s.d_force, s.h_force.data(), 4 * n * sizeof(float), hipMemcpyHostToDevice));
    }
    if (*ms < FLT_EPSILON) return 1;
    *maxRelError = 0.0f;
    if (crossCheck) {  // src/nbody.cu:406-418
        float mx = 0.0f;
        for (size_t i = 0; i < 4 * n; i++) {
            if ((i + 1) % 4 == 0) continue;
            const float err = relError(s.h_force[i], s.golden[i]);
            if (err > mx) mx = err;
        }
        *maxRelError = mx;
    }
    if (integrateGravitation_HIP(s.d_pm, s.d_vm, s.d_force, kTimestep, kDamping, n) <= 0.0f) {
        fprintf(stderr, "integrate failed: %s\n", nbody_hip_last_error());
        return 1;
    }
    return 0;
}

}  // namespace

int main(int argc, char** argv)
{
    int kParticles = kDefaultKParticles, maxIterations = 0, cycleAfter = 0;
    int idxFirstAlgorithm = 0, bPrintListOnly = 0, bUseGraphics = 0, bVerbose = 0;
    unsigned nSeed = kDefaultSeed;
    const char* ic = "reference";
    enum { OPT_IC = 1000, OPT_GOLDEN };
    static const struct option cli_options[] = {
        {"bodies", required_argument, nullptr, 'n'},     {"gpus", required_argument, nullptr, 'g'},
        {"no-cpu", no_argument, &g_bNoCPU, 1},           {"no-crosscheck", no_argument, &g_bCrossCheck, 0},
        {"iterations", required_argument, nullptr, 'i'}, {"cycle-after", required_argument, nullptr, 'c'},
        {"seed", required_argument, nullptr, 's'},       {"list", no_argument, nullptr, 'l'},
        {"algorithm", required_argument, nullptr, 'a'},  {"graphics", no_argument, nullptr, 'G'},
        {"verbose", no_argument, nullptr, 'v'},          {"help", no_argument, nullptr, 'h'},
        {"ic", required_argument, nullptr, OPT_IC},      {"golden", required_argument, nullptr, OPT_GOLDEN},
        {nullptr, 0, nullptr, 0}};

    if (hipGetDeviceCount(&g_maxGPUs) != hipSuccess) g_maxGPUs = 0;
    g_numGPUs = g_maxGPUs;

    while (1) {
        const int option = getopt_long(argc, argv, "n:s:i:c:g:la:Gvh", cli_options, nullptr);
        if (option == -1) break;
        switch (option) {
        case 0: break;
        case 'c': {
            int v;
            if (sscanf(optarg, "%d", &v) != 1) {
                fprintf(stderr, "ERROR: Couldn't parse integer argument for '--cycle-after'\n");
                return 1;
            }
            if (v < 1) {
                fprintf(stderr, "ERROR: Requested cycle size less than 1\n");
                return 1;
            }
            cycleAfter = v;
        } break;
        case 'i': {
            int v;
            if (sscanf(optarg, "%d", &v) != 1) {
                fprintf(stderr, "ERROR: Couldn't parse integer argument for '--iterations'\n");
                return 1;
            }
            if (v < 1) {
                fprintf(stderr, "ERROR: Requested number of iterations less than 1\n");
                return 1;
            }
            maxIterations = v;
        } break;
        case 'n': {
            int v;
            if (sscanf(optarg, "%d", &v) != 1) {
                fprintf(stderr, "ERROR: Couldn't parse integer argument for '--bodies'\n");
                return 1;
            }
            if (v < 1) {
                fprintf(stderr, "ERROR: Requested number of bodies less than 1\n");
                return 1;
            }
            kParticles = v;
        } break;
        case 's': {
            unsigned v;
            if (sscanf(optarg, "%u", &v) != 1) {
                fprintf(stderr, "ERROR: Couldn't parse integer argument for '--seed'\n");
                return 1;
            }
            if (v < 1) {
                fprintf(stderr, "ERROR: Seed must be nonzero integer value\n");
                return 1;
            }
            nSeed = v;
        } break;
        case 'g': {
            int v;
            if (sscanf(optarg, "%d", &v) != 1) {
                fprintf(stderr, "ERROR: Couldn't parse integer argument for '--gpus'\n");
                return 1;
            }
            if (g_maxGPUs < 1) {
                if (v == 0) {
                    g_numGPUs = 0;
                    break;
                }
                fprintf(stderr, "ERROR: No GPUs available, cannot handle '--gpus' argument.\n");
                return 1;
            }
            if (v < 1) {
                fprintf(stderr, "Requested number of GPUs less than 1, disabling GPU algorithms.\n");
                v = 0;
            }
            if (v > g_maxGPUs) {
                fprintf(stderr, "Requested %d GPUs, but only have %d, using all available GPUs.\n", v, g_maxGPUs);
                v = g_maxGPUs;
            }
            g_numGPUs = v;
        } break;
        case 'a': {
            const char* type;
            int v;
            idxFirstAlgorithm = -1;
            if (sscanf(optarg, "%d", &v) == 1) {
                type = "index";
                if (v >= 0 && v <= maxAlgorithmIdx()) idxFirstAlgorithm = v;
            } else {
                type = "name";
                for (v = 0; kAlgorithms[v].name; v++) {
                    if (strcasecmp(optarg, kAlgorithms[v].name) == 0) {
                        idxFirstAlgorithm = v;
                        break;
                    }
                    if (strncasecmp(optarg, kAlgorithms[v].name, strlen(optarg)) == 0) idxFirstAlgorithm = v;
                }
            }
            if (idxFirstAlgorithm == -1) {
                fprintf(stderr, "Invalid algorithm %s '%s'\n", type, optarg);
                print_algorithms();
                return 1;
            }
        } break;
        case 'G': bUseGraphics = 1; break;
        case 'l': bPrintListOnly = 1; break;
        case 'v': bVerbose = 1; break;
        case OPT_IC:
            if (strcmp(optarg, "reference") && strcmp(optarg, "ball") && strcmp(optarg, "plummer")) {
                fprintf(stderr, "ERROR: --ic must be reference, ball or plummer\n");
                return 1;
            }
            ic = optarg;
            break;
        case OPT_GOLDEN:
            if (strcmp(optarg, "fp64") && strcmp(optarg, "cpu_soa")) {
                fprintf(stderr, "ERROR: --golden must be fp64 or cpu_soa\n");
                return 1;
            }
            g_bGoldenSOA = !strcmp(optarg, "cpu_soa");
            break;
        case 'h':
        case '?': print_usage(); return 1;
        }
    }
    if (bUseGraphics) {
        fprintf(stderr, "Graphics requested, but OpenGL support not compiled in.\n");
        bUseGraphics = 0;
    }
    if (g_bNoCPU) g_bCrossCheck = 0;
    if (bPrintListOnly) {
        print_algorithms();
        return 0;
    }
    if (g_numGPUs > 1) {
        fprintf(stderr, "Using 1 of %d GPUs (this build computes on one).\n", g_numGPUs);
        g_numGPUs = 1;
    }

    State s;
    s.n = (size_t)kParticles * 1024;
    const size_t n = s.n;
    fprintf(stderr, "Running simulation with %u particles", (unsigned)n);
    if (g_bCrossCheck) fprintf(stderr, ", crosscheck enabled");
    if (g_bCrossCheck && g_bGoldenSOA) fprintf(stderr, " (CPU_SOA golden)");
    if (g_bNoCPU) fprintf(stderr, ", CPU disabled");
    if (g_numGPUs) fprintf(stderr, ", up to %u GPUs", g_numGPUs);
    fprintf(stderr, "\n");
    if (g_numGPUs < 1) {
        fprintf(stderr, "ERROR: no GPU available, and every algorithm of this build needs one\n");
        return 1;
    }

    std::vector<float> pm(4 * n), vm(4 * n);
    int icrc;
    if (!strcmp(ic, "ball"))
        icrc = nbody_uniform_ball(pm.data(), vm.data(), n, nSeed, (float)(1.54 * std::pow(n / 16384.0, 1.0 / 3.0)),
                                  1.0f, 1.0f);
    else if (!strcmp(ic, "plummer"))
        icrc = nbody_plummer(pm.data(), vm.data(), n, nSeed, 1.0f, 1.0f, 1);
    else
        icrc = nbody_random_unit_bodies(pm.data(), vm.data(), n, nSeed, kScale, kVelocityScale);
    if (icrc) {
        fprintf(stderr, "Failed to generate initial conditions.\n");
        return 1;
    }
    s.h_pm.resize(4 * n);
    s.h_force.resize(4 * n);
    s.golden.assign(4 * n, 0.0f);
    s.h_golden.resize(4 * n);
    HIP_OK(hipMalloc((void**)&s.d_pm, 4 * n * sizeof(float)));
    HIP_OK(hipMalloc((void**)&s.d_vm, 4 * n * sizeof(float)));
    HIP_OK(hipMalloc((void**)&s.d_force, 4 * n * sizeof(float)));
    HIP_OK(hipMalloc((void**)&s.d_golden, 4 * n * sizeof(double)));
    HIP_OK(hipMemcpy(s.d_pm, pm.data(), 4 * n * sizeof(float), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(s.d_vm, vm.data(), 4 * n * sizeof(float), hipMemcpyHostToDevice));

    int name_width = 0;
    for (int k = 0; kAlgorithms[k].name; k++) name_width = std::max(name_width, (int)strlen(kAlgorithms[k].name));

    // src/nbody.cu:1008-1089, without the keyboard and render hooks
    int algorithm_idx = idxFirstAlgorithm, steps = 0, iterations = 0, bStop = 0;
    auto start = std::chrono::steady_clock::now();
    while (!bStop) {
        float ms = 0.0f, err = 0.0f;
        const AlgoDef& algorithm = kAlgorithms[algorithm_idx];
        const auto now = std::chrono::steady_clock::now();
        if (ComputeGravitation(s, &ms, &err, algorithm, g_bCrossCheck)) goto next_algorithm;
        if (bVerbose || std::chrono::duration_cast<std::chrono::milliseconds>(now - start).count() >= 100) {
            const double interactionsPerSecond = (double)n * n * 1000.0f / ms;
            const double flops = interactionsPerSecond * (3 + 6 + 4 + 1 + 6) * 1e-3;
            start = now;
            if (interactionsPerSecond > 1e9)
                fprintf(stdout, "\r%*s: %8.2f ms = %8.3fx10^9 interactions/s (%9.2lf GFLOPS)", name_width,
                        algorithm.name, ms, interactionsPerSecond / 1e9, flops * 1e-6);
            else
                fprintf(stdout, "\r%*s: %8.2f ms = %8.3fx10^6 interactions/s (%9.2lf GFLOPS)", name_width,
                        algorithm.name, ms, interactionsPerSecond / 1e6, flops * 1e-6);
            if (g_bCrossCheck) fprintf(stdout, " (Rel. error: %E)\n", err);
            else fprintf(stdout, "\n");
            fflush(stdout);
        }
        steps++;
        if (cycleAfter && steps % cycleAfter == 0) {
        next_algorithm:
            steps = 0;
            algorithm_idx++;
            if (!kAlgorithms[algorithm_idx].name) {
                algorithm_idx = 0;
                iterations++;
            }
        } else if (!cycleAfter) {
            iterations++;
        }
        if (maxIterations && iterations >= maxIterations) bStop = 1;
    }
    (void)hipFree(s.d_pm);
    (void)hipFree(s.d_vm);
    (void)hipFree(s.d_force);
    (void)hipFree(s.d_golden);
    return 0;
}
