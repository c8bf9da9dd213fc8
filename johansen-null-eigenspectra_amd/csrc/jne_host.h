// jne_host.h -- host-side pieces shared by the storage code, the simulation
// driver and the command-line tool (not part of the C ABI).
#pragma once

#include <stddef.h>
#include <stdint.h>

#include <chrono>
#include <string>
#include <vector>

namespace jne_host {

// Thread-local message of the last failed host call (jne_host_last_error).
void set_error(const std::string& msg);
const std::string& last_error();

// "1,234,567" (src/display_utils.rs:117-130)
std::string commas(uint64_t n);
// "1 hour 2 minutes 5.00 seconds" (src/display_utils.rs:54-100)
std::string format_duration(double seconds);
// "estimated remaining: ...", "unknown", "completing soon" (src/display_utils.rs:151-170)
std::string format_remaining_time(double elapsed_s, uint64_t completed, uint64_t total);
// the percentile table of src/display_utils.rs:235-279
void print_percentiles_table(const std::string& model_name, const std::string& statistic, const double* p,
                             const double* v, size_t n, uint64_t total_count);

// "Model m: <description>" (src/johansen_models.rs:53-63,137-141)
const char* model_description(int model);
std::string model_display(int model);

// data/eigenvalues_model{m}_dim{d}_steps{T}.dat (src/data_storage/simulation.rs:96-122);
// creates data_dir when missing.
std::string dat_filename(const std::string& data_dir, int model, int dim, uint64_t steps, bool create_dir);

// A whole .dat file in memory (reader.rs semantics).
struct DatContents {
  uint8_t model = 0, dim = 0;
  uint32_t steps = 0;
  int nf = 0;                   // eigenvalues per record (0 when no records)
  bool has_trailer = false;     // ends with EOF_MARK + count + nF
  uint64_t data_end = 18;       // byte offset just past the last whole record
  std::vector<uint32_t> seeds;  // records in file order
  std::vector<double> eigs;     // seeds.size() * nf
};

// Reads path like reader.rs:23-216.  Returns 0, or JNE_ERR_IO /
// JNE_ERR_FORMAT with the message in last_error().  want_data = false skips
// storing the eigenvalues (seeds and counts only); force_scan ignores a
// trailer and scans whole records.
int read_dat(const char* path, bool want_data, DatContents* out, bool force_scan = false);

// Percentile by rank p*(n-1) with linear interpolation on ascending values
// (src/simulation_analyzers.rs:4-18).
double percentile_sorted(const double* sorted, size_t n, double p);

using Clock = std::chrono::steady_clock;

}  // namespace jne_host
