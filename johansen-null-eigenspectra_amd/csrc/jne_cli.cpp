// jne_cli.cpp -- the command-line tool: same flags, defaults, messages and
// file names as the reference binary (src/cli.rs:13-329, src/main.rs:44-230),
// with the walks computed on the GPU through the C ABI.
//
// Added flags (no counterpart in the reference): --precision <name> picks the
// moment accumulation (fp64 default; auto, mfma_fp16x2, mfma_bf16x3,
// mfma_bf16x2, mfma_fp32, fp32_kahan) and --device <int> the HIP device.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <iostream>
#include <string>
#include <thread>
#include <vector>

#include "../../include/jne.h"
#include "jne_host.h"

using namespace jne_host;

namespace {

struct CliArgs {
  long long threads = -1;  // --threads
  uint64_t steps = 10000;
  uint64_t runs = 10000000;
  uint64_t dim_start = 1, dim_end = 12;
  std::vector<int> models;  // empty: all
  bool quiet = false;
  int precision = JNE_PREC_FP64;
  int device = 0;
};

unsigned logical_cores() {
  const unsigned n = std::thread::hardware_concurrency();
  return n ? n : 1;
}

void print_help(const char* prog) {
  printf("Usage: %s [OPTIONS]\n\n", prog);
  printf("Options:\n");
  printf("  --threads <int>      number of threads for parallel computation (default: %u logical cores)\n",
         logical_cores());
  printf("  --steps <int>        number of simulation steps (default: 10,000)\n");
  printf("  --runs <int>         number of runs per model (default: 10,000,000)\n");
  printf("  --dim-start <int>    starting matrix dimension (default: 1)\n");
  printf("  --dim-end <int>      ending matrix dimension (default: 12)\n");
  printf("  --dim <int>          run a single dimension (sets start and end to the same value)\n");
  printf("  --model <list>       comma separated list of model numbers to compute (default: 0,1,2,3,4)\n");
  printf("  --quiet              suppress progress output\n");
  printf("  --precision <name>   moment accumulation: fp64 (default), auto, mfma_fp16x2, mfma_bf16x3, "
         "mfma_bf16x2, mfma_fp32, fp32_kahan\n");
  printf("  --device <int>       HIP device (default: 0)\n");
  printf("  -h, --help           show this help message\n");
  printf("  -v, --version        show version information\n\n");
  printf("Model Numbers:\n");
  for (int m = 0; m < 5; ++m) printf("  %d: %s\n", m, model_description(m));
  printf("\nExamples:\n");
  printf("  %s --threads 4 --steps 5,000 --runs 1,000,000\n", prog);
  printf("  %s --dim 5 --threads 8\n", prog);
  printf("  %s --dim-start 2 --dim-end 8 --runs 500,000\n", prog);
  printf("  %s --model 0,2 --runs 100,000\n", prog);
}

// a positive integer with optional commas (cli.rs:142-163)
bool parse_number(int argc, char** argv, int i, const char* name, uint64_t* out) {
  if (i + 1 >= argc) {
    fprintf(stderr, "Error: %s parameter requires a value\n", name);
    return false;
  }
  std::string s;
  for (const char* p = argv[i + 1]; *p; ++p)
    if (*p != ',') s.push_back(*p);
  bool ok = !s.empty() && s.size() <= 19;
  for (char c : s) ok = ok && c >= '0' && c <= '9';
  if (!ok) {
    fprintf(stderr, "Error: %s parameter must be a positive integer (commas allowed, e.g., 1,000,000)\n", name);
    return false;
  }
  *out = strtoull(s.c_str(), nullptr, 10);
  return true;
}

bool parse_models(const std::string& s, std::vector<int>* out, std::string* err) {  // cli.rs:178-199
  size_t start = 0;
  while (start <= s.size()) {
    size_t comma = s.find(',', start);
    if (comma == std::string::npos) comma = s.size();
    std::string part = s.substr(start, comma - start);
    const size_t a = part.find_first_not_of(" \t"), b = part.find_last_not_of(" \t");
    part = a == std::string::npos ? "" : part.substr(a, b - a + 1);
    if (!part.empty()) {
      bool digits = part.size() <= 3;
      for (char c : part) digits = digits && c >= '0' && c <= '9';
      const int v = digits ? atoi(part.c_str()) : -1;
      if (v < 0 || v > 4) {
        *err = "無效的模型代號: " + part;
        return false;
      }
      out->push_back(v);
    }
    start = comma + 1;
  }
  if (out->empty()) {
    *err = "模型列表不可為空";
    return false;
  }
  return true;
}

bool validate_threads(long long threads) {  // cli.rs:236-270
  const unsigned avail = logical_cores();
  if ((unsigned long long)threads > avail) {
    fprintf(stderr, "Warning: specified thread count (%lld) exceeds available logical cores (%u)\n", threads, avail);
    fprintf(stderr,
            "This may cause performance degradation due to thread contention. Recommend using no more than %u "
            "threads.\n",
            avail);
    if ((unsigned long long)threads > 2ull * avail) {
      fprintf(stderr, "Error: excessive thread count rejected to avoid system overload.\n");
      return false;
    }
    printf("Do you want to continue? (y/N): ");
    fflush(stdout);
    std::string input;
    std::getline(std::cin, input);
    const size_t a = input.find_first_not_of(" \t\r\n"), b = input.find_last_not_of(" \t\r\n");
    input = a == std::string::npos ? "" : input.substr(a, b - a + 1);
    for (char& c : input) c = (char)tolower(c);
    if (input != "y" && input != "yes") {
      printf("Execution cancelled.\n");
      return false;
    }
  }
  return true;
}

// 0 ok, 1 exit cleanly (help, version, error already printed)
int parse(int argc, char** argv, CliArgs* a) {
  if (argc > 1 && (!strcmp(argv[1], "--help") || !strcmp(argv[1], "-h"))) {
    print_help(argv[0]);
    return 1;
  }
  if (argc > 1 && (!strcmp(argv[1], "--version") || !strcmp(argv[1], "-v"))) {
    printf("johansen-null-eigenspectra version 0.8.0 (MI355X engine)\n");
    return 1;
  }
  for (int i = 1; i < argc;) {
    const std::string arg = argv[i];
    uint64_t v = 0;
    if (arg == "--threads") {
      if (!parse_number(argc, argv, i, "--threads", &v)) return 1;
      a->threads = (long long)v;
      i += 2;
    } else if (arg == "--steps") {
      if (!parse_number(argc, argv, i, "--steps", &v)) return 1;
      a->steps = v;
      i += 2;
    } else if (arg == "--runs") {
      if (!parse_number(argc, argv, i, "--runs", &v)) return 1;
      a->runs = v;
      i += 2;
    } else if (arg == "--dim-start") {
      if (!parse_number(argc, argv, i, "--dim-start", &v)) return 1;
      a->dim_start = v;
      i += 2;
    } else if (arg == "--dim-end") {
      if (!parse_number(argc, argv, i, "--dim-end", &v)) return 1;
      a->dim_end = v;
      i += 2;
    } else if (arg == "--dim") {
      if (!parse_number(argc, argv, i, "--dim", &v)) return 1;
      a->dim_start = a->dim_end = v;
      i += 2;
    } else if (arg == "--model") {
      if (i + 1 >= argc) {
        fprintf(stderr, "Error: --model parameter requires a value\n");
        return 1;
      }
      std::string err;
      a->models.clear();
      if (!parse_models(argv[i + 1], &a->models, &err)) {
        fprintf(stderr, "Error: %s\n", err.c_str());
        return 1;
      }
      i += 2;
    } else if (arg == "--quiet") {
      a->quiet = true;
      i += 1;
    } else if (arg == "--precision") {
      if (i + 1 >= argc) {
        fprintf(stderr, "Error: --precision parameter requires a value\n");
        return 1;
      }
      const std::string p = argv[i + 1];
      if (p == "fp64") a->precision = JNE_PREC_FP64;
      else if (p == "mfma_fp16x2") a->precision = JNE_PREC_MFMA_FP16X2;
      else if (p == "mfma_bf16x3") a->precision = JNE_PREC_MFMA_BF16X3;
      else if (p == "mfma_bf16x2") a->precision = JNE_PREC_MFMA_BF16X2;
      else if (p == "mfma_fp32") a->precision = JNE_PREC_MFMA_FP32;
      else if (p == "fp32_kahan") a->precision = JNE_PREC_FP32_KAHAN;
      else if (p == "auto") a->precision = JNE_PREC_AUTO;
      else {
        fprintf(stderr, "Error: unknown precision '%s'\n", p.c_str());
        return 1;
      }
      i += 2;
    } else if (arg == "--device") {
      if (!parse_number(argc, argv, i, "--device", &v)) return 1;
      a->device = (int)v;
      i += 2;
    } else {
      fprintf(stderr, "Error: unknown argument '%s'\n", arg.c_str());
      fprintf(stderr, "Use --help to see available options\n");
      return 1;
    }
  }
  // cli.rs:202-226
  if (a->dim_start > a->dim_end) {
    fprintf(stderr, "Error: dimension start (%llu) cannot be greater than end (%llu)\n",
            (unsigned long long)a->dim_start, (unsigned long long)a->dim_end);
    return 1;
  }
  if (a->steps == 0 || a->runs == 0 || a->dim_start == 0) {
    fprintf(stderr, "Error: steps, runs, and dimension parameters must be greater than 0\n");
    return 1;
  }
  // the engine's limits (the reference takes any usize; its file stores steps as u32)
  if (a->dim_end > JNE_MAX_DIM) {
    fprintf(stderr, "Error: dimension %llu is above the supported maximum of %d\n", (unsigned long long)a->dim_end,
            JNE_MAX_DIM);
    return 1;
  }
  if (a->steps < 2 || a->steps > (1ull << 30)) {
    fprintf(stderr, "Error: steps must be between 2 and %s\n", commas(1ull << 30).c_str());
    return 1;
  }
  if (a->runs > 0xFFFFFFFFull) {
    fprintf(stderr, "Error: runs must be at most %s (seeds are u32)\n", commas(0xFFFFFFFFull).c_str());
    return 1;
  }
  if (a->threads >= 0 && !validate_threads(a->threads)) return 1;
  return 0;
}

std::string fixed6(double x) {
  char b[64];
  snprintf(b, sizeof b, "%.6f", x);
  return b;
}

// records of the configuration, filtered to seed <= runs; false when the
// count differs from runs (simulation.rs:57-89)
bool read_data(const std::string& path, uint64_t runs, std::vector<uint32_t>* seeds, std::vector<double>* eigs,
               int* nf, std::string* err) {
  uint64_t n = 0;
  int k = 0;
  if (jne_dat_read_info(path.c_str(), nullptr, nullptr, nullptr, &n, &k, nullptr) != 0) {
    *err = jne_host_last_error();
    return false;
  }
  std::vector<uint32_t> s(n);
  std::vector<double> e(n * (size_t)std::max(k, 0));
  uint64_t got = 0;
  if (jne_dat_read(path.c_str(), s.data(), e.data(), n, &got, &k) != 0) {
    *err = jne_host_last_error();
    return false;
  }
  seeds->clear();
  eigs->clear();
  for (uint64_t i = 0; i < got; ++i) {
    if (s[i] > runs) continue;
    seeds->push_back(s[i]);
    eigs->insert(eigs->end(), e.begin() + i * k, e.begin() + (i + 1) * k);
  }
  *nf = k;
  if (seeds->size() != runs) {
    const std::string f = std::to_string(seeds->size());
    *err = "Data count mismatch: expected " + std::to_string(runs) + " records, found " + f +
           " records. This indicates incomplete simulation data.\nSolutions:\n1. Use `read_all_data()` to read "
           "available data (" + f + " records)\n2. Run `run_simulation()` to complete the simulation\n3. Create a "
           "new EigenvalueSimulation with num_runs=" + f + " to match existing data";
    return false;
  }
  return true;
}

void show_statistics(const std::string& path, int model, uint64_t runs) {  // main.rs:19-41
  static const double kP[] = {0.5, 0.75, 0.8, 0.85, 0.9, 0.95, 0.975, 0.99};
  std::vector<uint32_t> seeds;
  std::vector<double> eigs;
  int nf = 0;
  std::string err;
  if (!read_data(path, runs, &seeds, &eigs, &nf, &err) || seeds.empty()) return;
  for (int stat = 0; stat < 2; ++stat) {
    std::vector<double> v(seeds.size());
    for (size_t i = 0; i < seeds.size(); ++i) {
      double a = stat == 0 ? 0.0 : -1.7976931348623157e308;
      for (int j = 0; j < nf; ++j) a = stat == 0 ? a + eigs[i * nf + j] : std::max(a, eigs[i * nf + j]);
      v[i] = a;
    }
    double out[8];
    if (jne_percentiles(v.data(), v.size(), kP, 8, out) != 0) return;
    print_percentiles_table(model_display(model), stat == 0 ? "Trace" : "MaxEig", kP, out, 8, seeds.size());
    if (stat == 0) printf("\n");
  }
}

}  // namespace

int main(int argc, char** argv) {
  CliArgs a;
  if (parse(argc, argv, &a)) return 0;
  const bool q = a.quiet;
  const unsigned host_threads = a.threads > 0 ? (unsigned)a.threads : logical_cores();
  if (!q) {
    if (a.threads > 0)
      printf("Using %llu threads for parallel computation\n", (unsigned long long)a.threads);
    else
      printf("Using default thread count: %u\n", host_threads);
  }
  std::vector<int> models = a.models;
  if (models.empty()) models = {0, 1, 2, 3, 4};

  jne_ctx* ctx = nullptr;
  if (jne_create(a.device, &ctx) != JNE_OK) {
    fprintf(stderr, "Error: %s\n", jne_last_error(nullptr));
    return 1;
  }
  if (!q) {
    printf("=== Large-scale Simulation Demo ===\n");
    printf("Starting simulation (supports resuming from checkpoint)...\n");
    printf("Configuration:\n");
    printf("  Steps: %s\n", commas(a.steps).c_str());
    printf("  Runs: %s\n", commas(a.runs).c_str());
    printf("  Dimension range: %llu - %llu\n", (unsigned long long)a.dim_start, (unsigned long long)a.dim_end);
    printf("  Threads: %u\n\n", host_threads);
  }
  int status = 0;
  for (uint64_t dim = a.dim_start; dim <= a.dim_end && status == 0; ++dim) {
    const auto t0 = Clock::now();
    if (!q) {
      printf("Simulation config: %llu dimensions, %s steps, %s runs\n", (unsigned long long)dim,
             commas(a.steps).c_str(), commas(a.runs).c_str());
      printf("Starting large-scale eigenvalue simulation (supports resuming from checkpoint)...\n");
      printf("Dimensions: %s, Steps: %s, Runs: %s\n", commas(dim).c_str(), commas(a.steps).c_str(),
             commas(a.runs).c_str());
    }
    fflush(stdout);
    bool fuse = q && models.size() > 1 && models.size() <= 5;
    for (size_t i = 0; fuse && i < models.size(); ++i)
      for (size_t j = i + 1; j < models.size(); ++j)
        if (models[i] == models[j]) fuse = false;
    if (fuse) {
      // quiet: the models of this dim share one moment pass per walk
      // (jne_run_simulation_models); the files are those of the loop below
      std::vector<std::string> paths;
      std::vector<const char*> cpaths;
      for (int model : models) paths.push_back(dat_filename("data", model, (int)dim, a.steps, true));
      for (const std::string& p : paths) cpaths.push_back(p.c_str());
      const int rc = jne_run_simulation_models(ctx, models.data(), (int)models.size(), (int)dim, (uint32_t)a.steps,
                                               a.runs, a.precision, cpaths.data(), 1);
      if (rc != JNE_OK) {
        fprintf(stderr, "Error: %s (dim %llu): %s\n", rc == JNE_ERR_NUMERIC ? "numeric failure" : "failed",
                (unsigned long long)dim, jne_host_last_error());
        status = 1;
      }
      continue;
    }
    for (int model : models) {
      const std::string path = dat_filename("data", model, (int)dim, a.steps, true);
      const int rc = jne_run_simulation(ctx, model, (int)dim, (uint32_t)a.steps, a.runs, a.precision, path.c_str(),
                                        q ? 1 : 0);
      if (rc != JNE_OK) {
        // the reference panics here; report and stop with a non-zero status
        fprintf(stderr, "Error: %s (model %d, dim %llu): %s\n", rc == JNE_ERR_NUMERIC ? "numeric failure" : "failed",
                model, (unsigned long long)dim, jne_host_last_error());
        status = 1;
        break;
      }
      if (!q) show_statistics(path, model, a.runs);
    }
    if (!q && status == 0) {
      const double s = std::chrono::duration<double>(Clock::now() - t0).count();
      printf("Simulation completed! Duration: %s\n", format_duration(s).c_str());
    }
  }
  jne_destroy(ctx);
  if (status) return status;

  if (!q) {  // main.rs:137-228
    printf("\n=== Result Reading Demo ===\n");
    printf("Starting to read model data...\n");
    const int model = models[0];
    printf("Simulation config: %llu dimensions, %s steps, %s runs\n", (unsigned long long)a.dim_start,
           commas(a.steps).c_str(), commas(a.runs).c_str());
    const std::string path = dat_filename("data", model, (int)a.dim_start, a.steps, true);
    std::vector<uint32_t> seeds;
    std::vector<double> eigs;
    int nf = 0;
    std::string err;
    if (read_data(path, a.runs, &seeds, &eigs, &nf, &err)) {
      printf("Successfully read data for model %s: %s records\n", model_display(model).c_str(),
             commas(seeds.size()).c_str());
      printf("First 5 data records as examples:\n");
      const size_t k = std::min<size_t>(5, seeds.size());
      size_t w_seed = 0, w_sum = 0, w_eig = 0;
      for (size_t i = 0; i < k; ++i) {
        double sum = 0.0;
        for (int j = 0; j < nf; ++j) {
          sum += eigs[i * nf + j];
          w_eig = std::max(w_eig, fixed6(eigs[i * nf + j]).size());
        }
        w_seed = std::max(w_seed, std::to_string(seeds[i]).size());
        w_sum = std::max(w_sum, fixed6(sum).size());
      }
      if (k == 0) w_seed = 3, w_sum = 8, w_eig = 8;
      std::vector<std::string> strs(k);
      size_t w_str = k ? 0 : 20;
      for (size_t i = 0; i < k; ++i) {
        for (int j = 0; j < nf; ++j) {
          std::string v = fixed6(eigs[i * nf + j]);
          if (v.size() < w_eig) v = std::string(w_eig - v.size(), ' ') + v;
          strs[i] += (j ? ", " : "") + v;
        }
        w_str = std::max(w_str, strs[i].size());
      }
      for (size_t i = 0; i < k; ++i) {
        double sum = 0.0;
        for (int j = 0; j < nf; ++j) sum += eigs[i * nf + j];
        printf("  Record %2zu: seed=%*u, eigenvalue sum=%*.6f, eigenvalues=[%-*s]\n", i + 1, (int)w_seed, seeds[i],
               (int)w_sum, sum, (int)w_str, strs[i].c_str());
      }
    } else {
      printf("Failed to read data: %s\n", err.c_str());
    }
    printf("\n=== All Models Status ===\n");
    for (int m : models) {
      const std::string p = dat_filename("data", m, (int)a.dim_start, a.steps, true);
      if (read_data(p, a.runs, &seeds, &eigs, &nf, &err))
        printf("  %s: %s data records\n", model_display(m).c_str(), commas(seeds.size()).c_str());
      else
        printf("  %s: No data or read failed\n", model_display(m).c_str());
    }
    printf("\nDemo completed!\n");
  }
  return 0;
}
