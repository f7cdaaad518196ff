// spk_driver.cpp -- the `spooky` executable: runs a SPOOKY problem directory
// (spooky.cfg + data/snapNNNN.h5) on one MI355X through libspk.
//
// Follows the reference's Supervisor (supervisor.cu:20-213) and InputOutput:
//   startup: Parameters from spooky.cfg + command line (spk_run.cpp);
//   initial state: a snapshot (--restart, hdf5_io.cu:110-252) or the
//     configuration's mean field; at t = 0 the initial outputs (check_output.cu);
//   main loop: CleanFieldDivergence + HydroMHDAdvance per step, then
//     CheckOutput: timevar rows (spooky_output_timevar.cu) every timevar_step,
//     snapshots (hdf5_io.cu:19-106) at the snapshot clock kept on the device.
#include <sys/stat.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "spk.h"
#include "spk_config.hpp"
#include "spk_io.h"
#include "spk_run.hpp"

using namespace spk;

namespace {

#define CALL(x)                                                                                   \
    do {                                                                                          \
        const int _e = (x);                                                                       \
        if (_e != SPK_OK) {                                                                       \
            std::fprintf(stderr, "spooky: %s failed (%d): %s\n", #x, _e, spk_last_error());       \
            std::exit(1);                                                                         \
        }                                                                                         \
    } while (0)
#define IOCALL(x)                                                                                 \
    do {                                                                                          \
        if ((x) != 0) {                                                                           \
            std::fprintf(stderr, "spooky: %s failed: %s\n", #x, spkio_last_error());              \
            std::exit(1);                                                                         \
        }                                                                                         \
    } while (0)

struct Run {
    RunParams rp;
    spk_ctx ctx = nullptr;
    spk_solver sol = nullptr;
    spk_fft_plan plan = nullptr;  // forward transform of a restart snapshot
    int vars[10];
    size_t ntc = 0;
    int nf = 0;
    std::vector<std::string> names;
    void *d_fields = nullptr, *d_dfields = nullptr, *d_real = nullptr;
    double t_lastvar = 0.0;
    int num_save = 0;
    std::string data_dir;
    std::vector<double> host_real;  // padded real buffer on the host

    double tm[8];  // t, dt, tremap, dt_hyp, dt_par, dt_sts, t_lastsnap, step
    void read_time() { CALL(spk_solver_get_time(sol, tm)); }

    // padded [NX][NY][NZ+2] <-> unpadded [NX][NY][NZ]
    void unpad(const double *src, double *dst) const {
        const spk_params &p = rp.p;
        const size_t rz = 2 * (size_t)(p.nz / 2 + 1);
        for (size_t r = 0; r < (size_t)p.nx * p.ny; r++) std::memcpy(dst + r * p.nz, src + r * rz, sizeof(double) * p.nz);
    }
    void pad(const double *src, double *dst) const {
        const spk_params &p = rp.p;
        const size_t rz = 2 * (size_t)(p.nz / 2 + 1);
        for (size_t r = 0; r < (size_t)p.nx * p.ny; r++) {
            std::memcpy(dst + r * rz, src + r * p.nz, sizeof(double) * p.nz);
            dst[r * rz + p.nz] = dst[r * rz + p.nz + 1] = 0.0;
        }
    }

    // WriteDataFile (hdf5_io.cu:19-106)
    void write_snapshot() {
        const spk_params &p = rp.p;
        const size_t n = (size_t)p.nx * p.ny * p.nz;
        CALL(spk_snapshot_real(sol));  // unsheared in a shearing box
        host_real.resize(2 * ntc * nf);
        CALL(spk_memcpy_d2h(ctx, host_real.data(), d_real, sizeof(double) * host_real.size()));
        std::vector<std::vector<double>> f(nf, std::vector<double>(n));
        std::vector<const double *> ptrs(nf);
        std::vector<const char *> cn(nf);
        for (int i = 0; i < nf; i++) {
            unpad(host_real.data() + 2 * ntc * i, f[i].data());
            ptrs[i] = f[i].data();
            cn[i] = names[i].c_str();
        }
        char fn[32];
        std::snprintf(fn, sizeof fn, "snap%04d.h5", num_save);
        const std::string path = data_dir + "/" + fn;
        std::printf("-----------------Writing snap %04d at time_save: %.2e \n", num_save, tm[0]);
        IOCALL(spkio_write_snapshot(path.c_str(), p.nx, p.ny, p.nz, nf, cn.data(), ptrs.data(), p.t_initial, tm[0],
                                    tm[6], t_lastvar, (int)tm[7]));
        num_save++;
    }

    // ReadDataFile (hdf5_io.cu:110-252)
    void read_snapshot(int restart_num) {
        const spk_params &p = rp.p;
        const int last = spkio_last_snapshot(data_dir.c_str());
        if (last < 0) {
            std::fprintf(stderr, "spooky: no snapshot in %s to restart from\n", data_dir.c_str());
            std::exit(1);
        }
        int k = restart_num;
        if (restart_num > last) {
            std::printf("!!!! Error: restart file not found, restarting from last available data file...\n");
            k = last;
        } else if (restart_num == -1) {
            k = last;
        }
        char fn[32];
        std::snprintf(fn, sizeof fn, "snap%04d.h5", k);
        const std::string path = data_dir + "/" + fn;
        std::printf("Data file selected for restarting: \t%s\n", path.c_str());
        double t_save, t_lastsnap;
        int step;
        IOCALL(spkio_read_header(path.c_str(), &t_save, &t_lastsnap, &t_lastvar, &step));
        const size_t n = (size_t)p.nx * p.ny * p.nz;
        std::vector<double> f(n);
        host_real.assign(2 * ntc * nf, 0.0);
        for (int i = 0; i < nf; i++) {
            IOCALL(spkio_read_field(path.c_str(), names[i].c_str(), f.data(), n));
            pad(f.data(), host_real.data() + 2 * ntc * i);
        }
        CALL(spk_memcpy_h2d(ctx, d_real, host_real.data(), sizeof(double) * host_real.size()));
        CALL(spk_fft_plan_create(ctx, p.nx, p.ny, p.nz, SPK_F64, nf, &plan));
        CALL(spk_fft_r2c(plan, d_real, d_fields));
        CALL(spk_ctx_sync(ctx));
        CALL(spk_solver_set_time(sol, t_save, step, t_lastsnap));
        num_save = k + 1;
        std::printf("t_lastsnap: %.2e \t t_lastvar: %.2e \n", t_lastsnap, t_lastvar);
    }

    // initial state without a snapshot: zero fields plus the configured mean field
    void initial_state() {
        const spk_params &p = rp.p;
        if (rp.init_spatial_structure) {
            std::fprintf(stderr,
                         "spooky: init.spatial_structure is problem code (initflow.cu) in the reference; write it as "
                         "%s/snap0000.h5 with the initial-condition writer and run with -r 0\n",
                         data_dir.c_str());
            std::exit(1);
        }
        std::vector<double> c(2 * ntc * nf, 0.0);
        if (rp.init_mean_field && p.mhd) {
            const double ntot = (double)p.nx * p.ny * p.nz;
            const double b0[3] = {rp.bx0, rp.by0, rp.bz0};
            for (int q = 0; q < 3; q++) c[2 * ntc * (vars[3] + q)] = b0[q] * ntot;  // k = 0 mode
        }
        CALL(spk_memcpy_h2d(ctx, d_fields, c.data(), sizeof(double) * c.size()));
        CALL(spk_solver_set_time(sol, p.t_initial, 0, 0.0));
    }

    // WriteTimevarOutputHeader / WriteTimevarOutput (spooky_output_timevar.cu)
    void timevar_header() {
        FILE *f = std::fopen((data_dir + "/timevar.spooky").c_str(), "a");
        if (!f) return;
        std::fprintf(f, "## This file contains the time evolution of the following quantities: \n## \t");
        for (const auto &v : rp.timevar_vars) std::fprintf(f, "%s\t", v.c_str());
        std::fprintf(f, "\n");
        std::fclose(f);
    }
    void timevar_row() {
        FILE *f = std::fopen((data_dir + "/timevar.spooky").c_str(), "a");
        if (!f) return;
        for (const auto &v : rp.timevar_vars) {
            // quantities the library does not compute stay -1, as the reference
            // writes for a name it does not know
            double o = -1.0;
            if (v == "t") o = tm[0];
            else if (spk_timevar(sol, v.c_str(), &o) != SPK_OK) o = -1.0;
            std::fprintf(f, "%.8e\t", o);
        }
        std::fprintf(f, "\n");
        std::fclose(f);
    }

    // WriteSpectrumOutputHeader / WriteSpectrumOutput (spooky_output_spectrum.cu:22-480)
    void spectrum_header() {
        int nb = 0;
        double dk = 0.0;
        CALL(spk_spectrum_bins(sol, &nb, &dk));
        FILE *f = std::fopen((data_dir + "/spectrum.spooky").c_str(), "a");
        if (!f) return;
        std::fprintf(f, "## This file contains the 1d (shell-integrated) energy spectral densities of the following "
                        "quantities: \n## \t");
        for (const auto &v : rp.spectra_vars) std::fprintf(f, "%s\t", v.c_str());
        std::fprintf(f, "\nwavevector \t");
        for (int i = 0; i < nb; i++) std::fprintf(f, "%.8e\t", i * dk);
        std::fprintf(f, "\n");
        std::fclose(f);
    }
    void spectrum_rows() {
        if (rp.spectra_vars.empty()) return;
        int nb = 0;
        CALL(spk_spectrum_bins(sol, &nb, nullptr));
        std::vector<double> sp((size_t)nb);
        FILE *f = std::fopen((data_dir + "/spectrum.spooky").c_str(), "a");
        if (!f) return;
        for (const auto &v : rp.spectra_vars) {
            // a name the library does not compute gets no line, as in the reference
            if (spk_spectrum_named(sol, v.c_str(), sp.data(), nb, nullptr) != SPK_OK) continue;
            std::fprintf(f, "t\t%.8e\t%s\t", tm[0], v.c_str());
            for (double x : sp) std::fprintf(f, "%.8e\t", x);
            std::fprintf(f, "\n");
        }
        std::fclose(f);
    }
    // WriteUserTimevarOutputHeader / WriteUserTimevarOutput (user_output_timevar.cu:18-112)
    void user_header() {
        FILE *f = std::fopen((data_dir + "/user-timevar.spooky").c_str(), "a");
        if (!f) return;
        std::fprintf(f, "## This file contains the time evolution of the following quantities: \n## \t");
        for (const auto &v : rp.user_timevar_vars) std::fprintf(f, "%s\t", v.c_str());
        std::fprintf(f, "\n");
        std::fclose(f);
    }
    void user_row() {
        FILE *f = std::fopen((data_dir + "/user-timevar.spooky").c_str(), "a");
        if (!f) return;
        for (const auto &v : rp.user_timevar_vars) {
            double o = -1.0;
            if (spk_user_timevar(sol, v.c_str(), &o) != SPK_OK) o = -1.0;
            std::fprintf(f, "%.8e\t", o);
        }
        std::fprintf(f, "\n");
        std::fclose(f);
    }

    // CheckTimeSeries (check_output.cu:23-57); the rows read the real fields of
    // the state, made once here
    void check_time_series() {
        const double t = tm[0];
        if ((t - t_lastvar) >= rp.toutput_time || t == 0.0) {
            CALL(spk_clean_divergence(sol));
            std::printf("Saving output at t = %.6e \t and step n. %d \n", t, (int)tm[7]);
            const bool user = !rp.user_timevar_vars.empty();
            if (t == 0.0) {
                timevar_header();
                spectrum_header();
                if (user) user_header();
            }
            if (t != 0.0) t_lastvar += rp.toutput_time;
            CALL(spk_complex2real(sol));
            timevar_row();
            spectrum_rows();
            if (user) user_row();
        }
    }
    // CheckSnapshot (check_output.cu:53-94); due comes from the device snapshot
    // clock, which the main loop advances exactly as CheckSnapshot does
    void check_snapshot(bool due) {
        const double t = tm[0];
        if (due || t == 0.0) {
            CALL(spk_clean_divergence(sol));
            std::printf("Saving data snap at t= %.6e \t and step n. %d \n", t, (int)tm[7]);
            write_snapshot();
        }
    }
    void check_output(bool snapshot_due) {
        check_time_series();
        check_snapshot(snapshot_due);
    }
};

}  // namespace

int main(int argc, char **argv) {
    CliArgs cli;
    std::string err;
    if (!parse_cli(argc, argv, &cli, &err)) {
        std::fprintf(stderr, "%s\n%s", err.c_str(), cli_usage());
        return 1;
    }
    if (cli.help) {
        std::printf("%s", cli_usage());
        return 0;
    }
    Config cfg;
    const std::string cfgpath = cli.input_dir + "/spooky.cfg";
    if (!cfg.read_file(cfgpath)) {
        std::fprintf(stderr, "Error reading configuration file %s: %s\n", cfgpath.c_str(), cfg.error().c_str());
        return 1;
    }
    Run R;
    if (!read_run_params(cfg, cli, &R.rp, &err)) {
        std::fprintf(stderr, "spooky: %s\n", err.c_str());
        return 1;
    }
    std::vector<std::string> msgs;
    const int consistent = check_params(&R.rp, &msgs);
    for (const auto &m : msgs) std::printf("%s\n", m.c_str());
    if (cli.print_params) {
        std::printf("%s\n", params_json(R.rp).c_str());
        return consistent ? 0 : 1;
    }
    if (!R.rp.configname.empty()) std::printf("Using config file: %s at %s.\n", R.rp.configname.c_str(), cfgpath.c_str());
    if (!consistent) return 1;

    const spk_params &p = R.rp.p;
    R.data_dir = R.rp.output_dir + "/data";
    mkdir(R.rp.output_dir.c_str(), 0755);
    mkdir(R.data_dir.c_str(), 0755);

    CALL(spk_ctx_create(0, nullptr, &R.ctx));
    CALL(spk_solver_create(R.ctx, &p, R.rp.bulk, &R.sol));
    CALL(spk_solver_info(R.sol, R.vars, &R.ntc));
    R.nf = R.vars[9];
    R.names = field_names(p);
    CALL(spk_solver_buffers(R.sol, &R.d_fields, &R.d_dfields, &R.d_real));
    std::printf("grid %d x %d x %d, box %g x %g x %g, %d fields, %s bulk, output in %s\n", p.nx, p.ny, p.nz, p.lx,
                p.ly, p.lz, R.nf, R.rp.bulk == SPK_F32 ? "FP32" : "FP64", R.rp.output_dir.c_str());

    const bool restart = cli.restart_num != -2 || R.rp.restart;
    if (restart) R.read_snapshot(cli.restart_num == -2 ? -1 : cli.restart_num);
    else R.initial_state();
    R.read_time();
    // initialDataDump (supervisor.cu:215-239): all outputs for a fresh start,
    // the time series only when restarting from snapshot 0
    if (!restart) R.check_output(false);
    else if (cli.restart_num == 0) R.check_time_series();

    // executeMainLoop (supervisor.cu:169-213)
    const auto t0 = std::chrono::steady_clock::now();
    double elapsed = 0.0;
    long nsteps = 0;
    while (R.tm[0] < p.t_final && elapsed < R.rp.max_walltime_elapsed * 3600.0) {
        const double lastsnap = R.tm[6];
        // the first step starts fresh; later ones continue from the last step's
        // inverse x pass (the outputs between steps only read the state)
        CALL(nsteps == 0 ? spk_main_loop(R.sol, 1) : spk_main_loop_resume(R.sol, 1));
        R.read_time();
        nsteps++;
        R.check_output(R.tm[6] != lastsnap);
        elapsed = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (cli.stats > 0 && nsteps % cli.stats == 0)
            std::printf("step %d  t = %.6e  dt = %.3e  %.1f steps/s\n", (int)R.tm[7], R.tm[0], R.tm[1], nsteps / elapsed);
        if (elapsed >= R.rp.max_walltime_elapsed * 3600.0)
            std::printf("The maximum wallclock elapsed time was reached. Stopping now. You can later resume from the "
                        "last snapshot available. \n");
    }
    const double cells = (double)p.nx * p.ny * p.nz;
    std::printf("Finished: %ld steps in %.3f s (%.1f steps/s, %.3e cell updates/s), t = %.6e\n", nsteps, elapsed,
                nsteps / std::max(elapsed, 1e-9), nsteps * cells / std::max(elapsed, 1e-9), R.tm[0]);
    if (R.plan) spk_fft_plan_destroy(R.plan);
    spk_solver_destroy(R.sol);
    spk_ctx_destroy(R.ctx);
    return 0;
}
