// param.h -- param.in reader of the tensum driver.
//
// Same grammar, section order and validation as the reference's Reader and
// Parameter (src/reader.h:50-229, src/parameter.cc:16-875): whitespace, ','
// and ';' separate tokens; '//', '#', '%' start line comments; '/* */' block
// comments; sections grid, numeric, material, constants, initial_condition,
// boundary, exact_soln, integrals, output in that order.  Errors are returned,
// not turned into exit(0) as in the reference (src/mpi_utils.h:29-35).
#pragma once

#include <map>
#include <string>
#include <vector>

#include "../../include/tsm.h"

namespace tsm {

struct BcSpec {
  std::string type;                       // slip noslip farfield inlet outlet pressure periodic
  std::map<std::string, std::string> fn;  // variable -> expression
};

struct Settings {
  tsm_params p{};
  std::string time_mode, time_scheme;
  std::vector<int> sample_list;
  std::string rnd_file_loc;
  unsigned rnd_per_sample = 0;
  std::map<std::string, double> constants;
  std::map<std::string, std::string> ic;  // temperature xvelocity yvelocity zvelocity pressure
  std::string ic_library;
  std::map<int, BcSpec> bc;
  bool exact_available = false;
  std::map<std::string, std::string> exact;
  std::string write_format;
  unsigned write_frequency = 1;
  int n_time_stamps = 1;
  bool online_stat = false;
  std::vector<std::string> write_variables;
  // integrals/force blocks: name and boundary tags (src/parameter.cc:630-653)
  std::vector<std::pair<std::string, std::vector<int>>> forces;
  std::vector<int> write_surfaces;
  bool write_restart = false;
  unsigned restart_frequency = 0;
  bool write_log = true, write_soln = true, has_global = false, find_error = false;
};

// Reads `file`.  Returns false with `err` set on any grammar or validation error.
bool read_settings(const std::string& file, Settings& s, std::string& err);

}  // namespace tsm
