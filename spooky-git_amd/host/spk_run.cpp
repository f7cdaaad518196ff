// spk_run.cpp -- Parameters from spooky.cfg and the command line (spk_run.hpp).
#include "spk_run.hpp"

#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace spk {

const char *cli_usage() {
    return "usage: spooky [--input-dir DIR] [--output-dir DIR] [-r|--restart N] [--stats N]\n"
           "              [-t|--time HH MM SS] [-n|--ngrid NX NY NZ] [--with-2d]\n"
           "  --input-dir   directory of spooky.cfg (default ./)\n"
           "  --output-dir  output directory (overrides output.output_dir)\n"
           "  -r N          restart from data/snapNNNN.h5 (-1: the most recent)\n"
           "  --stats N     print stats every N steps\n"
           "  -t HH MM SS   maximum wall-clock time (95% of it is used)\n"
           "  -n NX NY NZ   override the grid size\n"
           "  --with-2d     force a 2D (x, y) problem\n"
           "  --fp32-bulk   transforms and real-space products in FP32 (state, RK3 and RKL2 stay FP64)\n"
           "  --print-params  print the resolved parameters as JSON and exit\n";
}

static bool to_int(const char *s, int *v) {
    char *e = nullptr;
    const long x = std::strtol(s, &e, 10);
    if (!s[0] || *e) return false;
    *v = (int)x;
    return true;
}

bool parse_cli(int argc, char **argv, CliArgs *a, std::string *err) {
    for (int i = 1; i < argc; i++) {
        const std::string o = argv[i];
        auto need = [&](int k) {
            if (i + k >= argc) {
                *err = "option " + o + " needs " + std::to_string(k) + " value(s)";
                return false;
            }
            return true;
        };
        auto ints = [&](int k, int *dst) {
            if (!need(k)) return false;
            for (int j = 0; j < k; j++)
                if (!to_int(argv[i + 1 + j], dst + j)) {
                    *err = "option " + o + ": integer expected, got '" + argv[i + 1 + j] + "'";
                    return false;
                }
            i += k;
            return true;
        };
        if (o == "-h" || o == "--help") {
            a->help = true;
        } else if (o == "--input-dir") {
            if (!need(1)) return false;
            a->input_dir = argv[++i];
        } else if (o == "--output-dir") {
            if (!need(1)) return false;
            a->output_dir = argv[++i];
            a->output_dir_override = true;
        } else if (o == "-r" || o == "--restart") {
            if (!ints(1, &a->restart_num)) return false;
        } else if (o == "--stats") {
            if (!ints(1, &a->stats)) return false;
        } else if (o == "-t" || o == "--time") {
            int hms[3];
            if (!ints(3, hms)) return false;
            // main.cu:96-105: hours + minutes/60 + seconds/3600, 95% of it
            a->max_hours = 0.95 * (hms[0] + hms[1] / 60.0 + hms[2] / 3600.0);
        } else if (o == "-n" || o == "--ngrid") {
            int n[3];
            if (!ints(3, n)) return false;
            a->nx = n[0];
            a->ny = n[1];
            a->nz = n[2];
        } else if (o == "--with-2d") {
            a->with2d = true;
        } else if (o == "--print-params") {
            a->print_params = true;
        } else if (o == "--fp32-bulk") {
            a->fp32_bulk = true;
        } else {
            *err = "unknown option " + o;
            return false;
        }
    }
    return true;
}

std::vector<std::string> field_names(const spk_params &p) {
    std::vector<std::string> n;
    if (p.incompressible) n = {"vx", "vy", "vz"};
    if (p.mhd) n.insert(n.end(), {"bx", "by", "bz"});
    if (p.boussinesq || p.heat_equation) n.push_back("th");
    return n;
}

bool read_run_params(const Config &c, const CliArgs &cli, RunParams *r, std::string *err) {
    spk_params &p = r->p;
    std::memset(&p, 0, sizeof p);
    c.lookup_string("configname", &r->configname);
    // modules (parameters.cu:77-167)
    if (!c.lookup_int("modules.debug", &r->debug)) r->debug = 0;
    if (!c.lookup_bool("modules.with2d", &p.with2d)) p.with2d = 0;
    if (cli.with2d) p.with2d = 1;
    if (!c.lookup_bool("modules.incompressible", &p.incompressible)) p.incompressible = 0;
    if (!c.lookup_bool("modules.heat_equation", &p.heat_equation)) p.heat_equation = 0;
    if (!c.lookup_bool("modules.boussinesq", &p.boussinesq)) p.boussinesq = 0;
    if (!c.lookup_bool("modules.stratification", &p.stratification)) p.stratification = 0;
    if (!c.lookup_int("modules.strat_direction", &p.strat_direction)) p.strat_direction = 2;
    if (!c.lookup_bool("modules.mhd", &p.mhd)) p.mhd = 0;
    if (!c.lookup_bool("modules.anisotropic_diffusion", &p.anisotropic_diffusion)) p.anisotropic_diffusion = 0;
    if (!c.lookup_bool("modules.supertimestepping", &p.supertimestepping)) p.supertimestepping = 0;
    r->sts_variables.clear();
    if (p.supertimestepping) {
        if (c.lookup_string("modules.sts_algorithm", &r->sts_algorithm)) {
            if (!c.lookup_strings("modules.sts_variables", &r->sts_variables) || r->sts_variables.empty()) {
                // no variables: every parabolic term is evolved as normal
                r->sts_variables.clear();
                p.supertimestepping = 0;
            }
        }
    }
    if (!c.lookup_bool("modules.shearing", &p.shearing)) p.shearing = 0;
    if (!c.lookup_bool("modules.rotating", &p.rotating)) p.rotating = 0;
    // physics (parameters.cu:171-256)
    if (!c.lookup_float("physics.boxsize.[0]", &p.lx)) p.lx = 1.0;
    if (!c.lookup_float("physics.boxsize.[1]", &p.ly)) p.ly = 1.0;
    if (!c.lookup_float("physics.boxsize.[2]", &p.lz)) p.lz = 1.0;
    if (!c.lookup_int("physics.gridsize.[0]", &p.nx)) p.nx = 32;
    if (cli.nx > 0) p.nx = cli.nx;
    if (!c.lookup_int("physics.gridsize.[1]", &p.ny)) p.ny = 32;
    if (cli.ny > 0) p.ny = cli.ny;
    if (!c.lookup_int("physics.gridsize.[2]", &p.nz)) p.nz = 32;
    if (cli.nz > 0) p.nz = cli.nz;
    if (!c.lookup_float("physics.reynolds", &r->reynolds)) r->reynolds = 1.0;
    p.nu = 1. / r->reynolds;
    if (!c.lookup_float("physics.reynolds_magnetic", &r->reynolds_m)) r->reynolds_m = 1.0;
    p.nu_m = 1. / r->reynolds_m;
    if (!c.lookup_float("physics.reynolds_thermic", &r->reynolds_th)) r->reynolds_th = 1.0;
    p.nu_th = 1. / r->reynolds_th;
    if (!c.lookup_float("physics.reynolds_anisotropic", &p.reynolds_ani)) p.reynolds_ani = 1.0;
    if (!c.lookup_float("physics.reynolds_Braginskii", &r->reynolds_B)) r->reynolds_B = 1.0;
    if (!c.lookup_float("physics.brunt_vaissala_squared", &p.N2)) p.N2 = 0.0;
    if (!c.lookup_float("physics.mti_frequency_squared", &p.OmegaT2)) p.OmegaT2 = 0.0;
    if (!c.lookup_float("physics.conductivity_alpha", &r->ALPHA)) r->ALPHA = 0.0;
    if (!c.lookup_float("physics.omega", &p.omega)) p.omega = 0.0;
    if (!c.lookup_float("physics.shear", &p.shear)) p.shear = 0.0;
    if (!c.lookup_float("physics.sound_speed", &r->cs)) r->cs = 1.0;
    // code (parameters.cu:286-340)
    if (!c.lookup_float("code.cfl", &p.cfl)) p.cfl = 1.5;
    if (!c.lookup_float("code.cfl_par", &p.cfl_par)) p.cfl_par = 1.0;
    if (!c.lookup_float("code.cfl_sts", &p.cfl_sts)) p.cfl_sts = 1.0;
    if (!c.lookup_int("code.safety_sts", &p.safety_sts)) p.safety_sts = 1;
    if (!c.lookup_float("code.safety_source", &p.safety_source)) p.safety_source = 0.2;
    if (!c.lookup_float("code.t_initial", &p.t_initial)) p.t_initial = 0.0;
    if (!c.lookup_float("code.t_final", &p.t_final)) p.t_final = 1.0;
    if (!c.lookup_float("code.max_walltime_elapsed", &r->max_walltime_elapsed)) {
        r->max_walltime_elapsed = 42;
        if (cli.max_hours > 0.0) r->max_walltime_elapsed = cli.max_hours;
    }
    r->bulk = cli.fp32_bulk ? SPK_F32 : SPK_F64;
    if (!c.lookup_int("code.interface_check", &r->interface_check)) r->interface_check = 5;
    if (!c.lookup_bool("code.interface_output_file", &r->interface_output_file)) r->interface_output_file = 0;
    if (!c.lookup_bool("code.force_symmetries", &r->force_symmetries)) r->force_symmetries = 0;
    if (!c.lookup_int("code.symmetries_step", &r->symmetries_step)) r->symmetries_step = 20;
    if (!c.lookup_bool("code.antialiasing", &p.antialiasing)) p.antialiasing = 1;
    if (!c.lookup_bool("code.restart", &r->restart)) {
        r->restart = 0;
        if (cli.restart_num > -1) r->restart = 1;
    }
    // output (parameters.cu:348-441)
    if (!c.lookup_float("output.timevar_step", &r->toutput_time)) r->toutput_time = 1.0;
    if (!c.lookup_float("output.snapshot_step", &p.toutput_flow)) p.toutput_flow = 1.0;
    if (!c.lookup_float("output.dump_step", &r->toutput_dump)) r->toutput_dump = 1.0;
    if (!c.lookup_bool("output.vorticity", &r->output_vorticity)) r->output_vorticity = 0;
    // the reference leaves output_dir empty without the key; the input directory here
    r->output_dir = cli.input_dir;
    c.lookup_string("output.output_dir", &r->output_dir);
    if (cli.output_dir_override) r->output_dir = cli.output_dir;
    c.lookup_strings("output.timevar_vars", &r->timevar_vars);
    c.lookup_strings("output.spectra_vars", &r->spectra_vars);
    c.lookup_strings("output.user_timevar_vars", &r->user_timevar_vars);
    if (!c.lookup_int("output.profile_dir", &r->profile_dir)) r->profile_dir = 0;
    // init (parameters.cu:449-503)
    if (!c.lookup_bool("init.vortex.enable", &r->init_vortex)) r->init_vortex = 0;
    if (!c.lookup_float("init.vortex.a", &r->vortex_a)) r->vortex_a = 1.0;
    if (!c.lookup_float("init.vortex.b", &r->vortex_b)) r->vortex_b = 2.0;
    if (!c.lookup_bool("init.spatial_structure", &r->init_spatial_structure)) r->init_spatial_structure = 0;
    if (!c.lookup_bool("init.large_scale_noise.enable", &r->init_large_scale_noise)) r->init_large_scale_noise = 0;
    if (!c.lookup_float("init.large_scale_noise.amplitude", &r->per_amplitude_large)) r->per_amplitude_large = 0.0;
    if (!c.lookup_float("init.large_scale_noise.cut_length", &r->noise_cut_length)) r->noise_cut_length = 0.0;
    if (!c.lookup_bool("init.large_scale_2D_noise.enable", &r->init_large_scale_2D_noise))
        r->init_large_scale_2D_noise = 0;
    if (!c.lookup_float("init.large_scale_2D_noise.amplitude", &r->per_amplitude_large_2D))
        r->per_amplitude_large_2D = 0.0;
    if (!c.lookup_float("init.large_scale_2D_noise.cut_length", &r->noise_cut_length_2D)) r->noise_cut_length_2D = 0.0;
    if (!c.lookup_bool("init.white_noise.enable", &r->init_white_noise)) r->init_white_noise = 0;
    if (!c.lookup_float("init.white_noise.amplitude", &r->per_amplitude_noise)) r->per_amplitude_noise = 0.0;
    if (!c.lookup_bool("init.mean_field.enable", &r->init_mean_field)) r->init_mean_field = 0;
    if (!c.lookup_float("init.mean_field.bx0", &r->bx0)) r->bx0 = 0.0;
    if (!c.lookup_float("init.mean_field.by0", &r->by0)) r->by0 = 0.0;
    if (!c.lookup_float("init.mean_field.bz0", &r->bz0)) r->bz0 = 0.0;
    if (!c.lookup_bool("init.dump", &r->init_dump)) r->init_dump = 0;
    if (!c.lookup_bool("init.bench", &r->init_bench)) r->init_bench = 0;

    // sts variables -> field mask (RKLegendre ctor, rkl.cu:40-57)
    p.sts_mask = 0;
    if (p.supertimestepping) {
        if (r->sts_algorithm != "rkl2") {
            *err = "sts_algorithm \"" + r->sts_algorithm + "\" is not available: this solver runs \"rkl2\"";
            return false;
        }
        const std::vector<std::string> names = field_names(p);
        for (const auto &v : r->sts_variables) {
            int f = -1;
            for (size_t k = 0; k < names.size(); k++)
                if (names[k] == v) f = (int)k;
            if (f < 0) {
                *err = "sts variable \"" + v + "\" is not a field of this configuration";
                return false;
            }
            p.sts_mask |= 1 << f;
        }
    }
    return true;
}

static std::string jstr(const std::string &s) {
    std::string o = "\"";
    for (char c : s) {
        if (c == '"' || c == '\\') o.push_back('\\');
        if (c == '\n') {
            o += "\\n";
            continue;
        }
        o.push_back(c);
    }
    return o + "\"";
}
static std::string jlist(const std::vector<std::string> &v) {
    std::string o = "[";
    for (size_t i = 0; i < v.size(); i++) o += (i ? ", " : "") + jstr(v[i]);
    return o + "]";
}

std::string params_json(const RunParams &r) {
    const spk_params &p = r.p;
    char b[2048];
    std::snprintf(b, sizeof b,
                  "\"nx\": %d, \"ny\": %d, \"nz\": %d, \"lx\": %.17g, \"ly\": %.17g, \"lz\": %.17g, "
                  "\"incompressible\": %d, \"heat_equation\": %d, \"boussinesq\": %d, \"stratification\": %d, "
                  "\"strat_direction\": %d, \"mhd\": %d, \"anisotropic_diffusion\": %d, \"supertimestepping\": %d, "
                  "\"shearing\": %d, \"rotating\": %d, \"with2d\": %d, \"antialiasing\": %d, \"sts_mask\": %d, "
                  "\"safety_sts\": %d, \"nu\": %.17g, \"nu_m\": %.17g, \"nu_th\": %.17g, \"reynolds_ani\": %.17g, "
                  "\"N2\": %.17g, \"OmegaT2\": %.17g, \"omega\": %.17g, \"shear\": %.17g, \"cfl\": %.17g, "
                  "\"cfl_par\": %.17g, \"cfl_sts\": %.17g, \"safety_source\": %.17g, \"t_initial\": %.17g, "
                  "\"t_final\": %.17g, \"toutput_flow\": %.17g, \"toutput_time\": %.17g, \"toutput_dump\": %.17g, "
                  "\"max_walltime_elapsed\": %.17g, \"debug\": %d, \"restart\": %d, \"init_spatial_structure\": %d, "
                  "\"init_mean_field\": %d, \"bx0\": %.17g, \"by0\": %.17g, \"bz0\": %.17g",
                  p.nx, p.ny, p.nz, p.lx, p.ly, p.lz, p.incompressible, p.heat_equation, p.boussinesq, p.stratification,
                  p.strat_direction, p.mhd, p.anisotropic_diffusion, p.supertimestepping, p.shearing, p.rotating,
                  p.with2d, p.antialiasing, p.sts_mask, p.safety_sts, p.nu, p.nu_m, p.nu_th, p.reynolds_ani, p.N2,
                  p.OmegaT2, p.omega, p.shear, p.cfl, p.cfl_par, p.cfl_sts, p.safety_source, p.t_initial, p.t_final,
                  p.toutput_flow, r.toutput_time, r.toutput_dump, r.max_walltime_elapsed, r.debug, r.restart,
                  r.init_spatial_structure, r.init_mean_field, r.bx0, r.by0, r.bz0);
    return std::string("{") + b + ", \"configname\": " + jstr(r.configname) + ", \"output_dir\": " +
           jstr(r.output_dir) + ", \"sts_algorithm\": " + jstr(r.sts_algorithm) + ", \"sts_variables\": " +
           jlist(r.sts_variables) + ", \"timevar_vars\": " + jlist(r.timevar_vars) + ", \"spectra_vars\": " +
           jlist(r.spectra_vars) + ", \"user_timevar_vars\": " + jlist(r.user_timevar_vars) +
           ", \"fields\": " + jlist(field_names(p)) + ", \"bulk\": " + jstr(r.bulk == SPK_F32 ? "f32" : "f64") +
           "}";
}

int check_params(RunParams *r, std::vector<std::string> *msgs) {
    spk_params &p = r->p;
    int ok = 1;
    auto bad = [&](const char *m) {
        ok = 0;
        msgs->push_back(std::string("Error: ") + m);
    };
    if (p.incompressible && p.heat_equation) bad("incompressible and heat_equation modules are mutually exclusive");
    if (!p.incompressible && !p.heat_equation) bad("you have to chose either incompressible or heat_equation module");
    if (p.boussinesq && !p.incompressible) bad("Boussinesq requires incompressible module");
    if (p.stratification && !p.boussinesq) bad("stratification requires Boussinesq module");
    if (p.mhd && !p.incompressible) bad("mhd requires incompressible module");
    if (p.anisotropic_diffusion && !(p.mhd && p.boussinesq)) bad("anisotropic_diffusion requires mhd and boussinesq module");
    if (p.supertimestepping && r->sts_variables.size() > field_names(p).size())
        bad("the number of sts variables is larger than the total number of variables.");
    if (p.nx < 2 || p.ny < 2 || p.nz < 2)
        msgs->push_back("Error: please select a number of grid points >= 2 in each direction.");
    // shearing: snapshots only when the domain is exactly periodic (parameters.cu:548-562)
    if (p.shearing) {
        const double tperiodic = p.ly / (p.shear * p.lx);
        if (p.toutput_flow < tperiodic) p.toutput_flow = tperiodic;
        else p.toutput_flow = (double)(int)(p.toutput_flow / tperiodic);
        msgs->push_back("Overwriting output for data snapshots because of shearing");
    }
    return ok;
}

}  // namespace spk
