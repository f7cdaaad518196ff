// spk_run.hpp -- run parameters of the driver: the reference's Parameters
// (parameters.cu:77-505) read from spooky.cfg with the same keys and
// defaults, the command line of main.cu:30-134, and checkParameters
// (parameters.cu:510-566).
#pragma once

#include <string>
#include <vector>

#include "spk.h"
#include "spk_config.hpp"

namespace spk {

// main.cu:30-134 (argparse options)
struct CliArgs {
    std::string input_dir = "./";
    std::string output_dir;
    bool output_dir_override = false;
    int restart_num = -2;   // -2: no --restart; -1: most recent snapshot; n: snapNNNN
    int stats = -1;         // print stats every n steps (-1: never)
    double max_hours = -1;  // --time HH MM SS, 95% of it
    int nx = -1, ny = -1, nz = -1;
    bool with2d = false;
    bool help = false;
    bool print_params = false;  // print the resolved parameters as JSON and exit (no GPU used)
    bool fp32_bulk = false;     // transforms and products in FP32, state and RK3/RKL2 in FP64
};
// Parses argv; returns false with a message on a bad command line.
bool parse_cli(int argc, char **argv, CliArgs *a, std::string *err);
const char *cli_usage();

struct RunParams {
    spk_params p;                 // what the solver needs
    std::string configname;
    int debug = 0;
    std::string sts_algorithm;
    std::vector<std::string> sts_variables;
    double reynolds = 1, reynolds_m = 1, reynolds_th = 1, reynolds_B = 1, ALPHA = 0, cs = 1;
    double max_walltime_elapsed = 42;  // hours
    int interface_check = 5, interface_output_file = 0, force_symmetries = 0, symmetries_step = 20, restart = 0;
    double toutput_time = 1.0, toutput_dump = 1.0;
    int output_vorticity = 0, profile_dir = 0;
    std::string output_dir;
    std::vector<std::string> timevar_vars, spectra_vars, user_timevar_vars;
    // init section (parameters.cu:449-503)
    int init_vortex = 0, init_spatial_structure = 0, init_large_scale_noise = 0, init_large_scale_2D_noise = 0;
    int init_white_noise = 0, init_mean_field = 0, init_dump = 0, init_bench = 0;
    double vortex_a = 0, vortex_b = 0, per_amplitude_large = 0, noise_cut_length = 0;
    double per_amplitude_large_2D = 0, noise_cut_length_2D = 0, per_amplitude_noise = 0;
    double bx0 = 0, by0 = 0, bz0 = 0;
    int bulk = SPK_F64;           // precision of the transforms and products (--fp32-bulk)
};

// Fills r from cfg and the command line overrides; returns false with a
// message for settings this solver cannot run (an sts algorithm other than rkl2).
bool read_run_params(const Config &cfg, const CliArgs &cli, RunParams *r, std::string *err);

// checkParameters (parameters.cu:510-566): 1 when consistent; messages go to
// msgs.  Also applies its shearing adjustment of the snapshot interval.
int check_params(RunParams *r, std::vector<std::string> *msgs);

// Field names in the order of popVariablesGrid (parameters.cu:568-640).
std::vector<std::string> field_names(const spk_params &p);

// The resolved parameters as one JSON object.
std::string params_json(const RunParams &r);

}  // namespace spk
