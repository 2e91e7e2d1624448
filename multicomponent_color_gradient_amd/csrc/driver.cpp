// driver.cpp -- mcg_jet: the round-jet run of the reference's main.cu:17-120
// on the C ABI.  Same loop order and cadence: every step Macros -> [radial
// statistics from NSTATS_START_STEP] -> collide-stream -> inlet -> outlet;
// an MLUPS window every 100 steps (utilities/countmlups.cuh:20-64); every
// NOUTPUT steps (including step 0) a TKE sample, the radial files once
// statistics have samples, and a VTI file; the radial files again at the end.
//
// usage: mcg_jet [--reference] [--nx N --ny N --nz N] [--radius R] [--velocity U]
//                [--Re Re] [--We We] [--steps N] [--noutput N] [--stats-start N]
//                [--stats-sample N] [--out DIR] [--no-vti] [--kernel fused|twopass]
//                [--storage fp32|fp16|bf16]
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/mcg.h"

namespace {

int die(const char *what, int rc) {
    std::fprintf(stderr, "mcg_jet: %s failed (%d): %s %s\n", what, rc, mcg_last_error(), mcg_io_last_error());
    return 1;
}

#define TRY(what, expr)                   \
    do {                                  \
        const int rc_ = (expr);           \
        if (rc_ != MCG_OK) return die(what, rc_); \
    } while (0)

}  // namespace

int main(int argc, char **argv) {
    mcg_case c;
    mcg_case_reference_jet(&c);
    std::string out = ".";
    bool reference = false;
    int kernel = MCG_KERNEL_FUSED;
    int storage = MCG_STORE_FP32;
    double R = -1, U = -1, Re = -1, We = -1;
    for (int a = 1; a < argc; ++a) {
        const std::string k = argv[a];
        auto val = [&]() -> const char * {
            if (a + 1 >= argc) {
                std::fprintf(stderr, "missing value for %s\n", k.c_str());
                std::exit(2);
            }
            return argv[++a];
        };
        if (k == "--reference") reference = true;
        else if (k == "--nx") c.NX = std::atoi(val());
        else if (k == "--ny") c.NY = std::atoi(val());
        else if (k == "--nz") c.NZ = std::atoi(val());
        else if (k == "--radius") R = std::atof(val());
        else if (k == "--velocity") U = std::atof(val());
        else if (k == "--Re") Re = std::atof(val());
        else if (k == "--We") We = std::atof(val());
        else if (k == "--steps") c.NSTEP = std::atoi(val());
        else if (k == "--noutput") c.NOUTPUT = std::atoi(val());
        else if (k == "--stats-start") c.NSTATS_START_STEP = std::atoi(val());
        else if (k == "--stats-sample") c.NSTATS_SAMPLE = std::atoi(val());
        else if (k == "--out") out = val();
        else if (k == "--no-vti") c.write_vti_output = 0;
        else if (k == "--kernel") kernel = std::string(val()) == "twopass" ? MCG_KERNEL_TWOPASS : MCG_KERNEL_FUSED;
        else if (k == "--storage") {
            const std::string v = val();
            if (v == "fp32") storage = MCG_STORE_FP32;
            else if (v == "fp16") storage = MCG_STORE_FP16;
            else if (v == "bf16") storage = MCG_STORE_BF16;
            else {
                std::fprintf(stderr, "--storage must be fp32, fp16 or bf16\n");
                return 2;
            }
        }
        else {
            std::fprintf(stderr, "unknown option %s\n", k.c_str());
            return 2;
        }
    }
    if (!reference) {
        // re-derive the dependent constants of constants.cuh:32-76 for the chosen grid and jet
        if (R > 0) c.jet_radius = R;
        if (U > 0) c.jet_velocity = U;
        if (Re > 0) c.Re = Re;
        if (We > 0) c.We = We;
        c.jet_x0 = (c.NX - 1) / 2.0;
        c.jet_z0 = (c.NZ - 1) / 2.0;
        c.bubble_x0 = (c.NX - 1) / 2.0;
        c.bubble_y0 = (c.NY - 1) / 2.0;
        c.bubble_z0 = (c.NZ - 1) / 2.0;
        const double nu = 2.0 * c.jet_radius * c.jet_velocity / c.Re;
        c.taur = c.taub = 0.5 + nu / (1.0 / 3.0);
        c.sigma = c.rhob0 * c.jet_velocity * c.jet_velocity * 2.0 * c.jet_radius / c.We;
    }
    if (c.NOUTPUT <= 0) c.NOUTPUT = 1;
    if (c.NSTEP <= c.NSTATS_START_STEP)
        std::printf("Warning: NSTEP (%d) <= NSTATS_START_STEP (%d), so radial statistics will not start in this run.\n",
                    c.NSTEP, c.NSTATS_START_STEP);

    char case_dir[4096], mean_dir[4096], vti_dir[4096];
    TRY("output directories", mcg_output_dirs(&c, out.c_str(), case_dir, mean_dir, vti_dir, sizeof(case_dir)));
    TRY("case_info.txt", mcg_write_case_info(&c, case_dir));
    TRY("tke files", mcg_tke_files_reset(mean_dir));

    mcg_solver *s = nullptr;
    TRY("create", mcg_create(&c, storage, nullptr, &s));
    TRY("kernel", mcg_set_kernel(s, kernel));
    TRY("init", mcg_init_jet(s));
    std::printf("%s | %dx%dx%d | %d steps | out %s\n", mcg_version(), c.NX, c.NY, c.NZ, c.NSTEP, case_dir);

    const double cells = double(c.NX) * c.NY * c.NZ;
    const int window = 100;
    double window_ms = 0.0, total_ms = 0.0;
    float tke_state[3] = {0.0f, 0.0f, 1.0f};
    bool stats_announced = false;
    for (int step = 0; step < c.NSTEP;) {
        // run up to the next step after which something is reported
        int last = step;
        while (last + 1 < c.NSTEP && (last + 1) % window != 0 && last % c.NOUTPUT != 0) ++last;
        const int n = last - step + 1;
        double ms = 0.0;
        TRY("step", mcg_time_steps(s, n, &ms));
        window_ms += ms;
        total_ms += ms;
        step = last + 1;  // steps 0..last done
        int32_t started = 0;
        int64_t start_step = 0, samples = 0;
        TRY("statistics state", mcg_stats_info(s, &started, &start_step, &samples));
        if (started && !stats_announced) {
            std::printf("Starting radial temporal statistics at step %lld | stats_start_step = %d\n", (long long)start_step,
                        c.NSTATS_START_STEP);
            stats_announced = true;
        }
        if (step % window == 0) {
            const double mlups = window * cells / (window_ms * 1.0e3);
            const double avg = step * cells / (total_ms * 1.0e3);
            std::printf("step %d/%d | dt_window = %.3f ms | MLUPS = %.3f | MLUPS(avg) = %.3f\n", step, c.NSTEP, window_ms,
                        mlups, avg);
            window_ms = 0.0;
        }
        if (last % c.NOUTPUT == 0) {
            TRY("health", mcg_check_health(s));
            double tke = 0.0;
            TRY("tke", mcg_tke(s, &tke));
            const float prev = tke_state[1];
            float vals[3];
            mcg_tke_update(tke_state, float(tke), last, c.NOUTPUT, vals);
            TRY("tke files", mcg_append_tke(mean_dir, vals));
            std::printf("step %d | tke_prev = %e | tke_avg = %e | delta_tke_avg = %e | %s\n", last, double(prev),
                        double(vals[1]), double(vals[2]), started ? "uy_radial_avg = on" : "uy_radial_avg = waiting");
            if (started && samples > 0) TRY("radial files", mcg_write_radial(s, mean_dir));
            if (c.write_vti_output) TRY("vti", mcg_write_vti_step(s, vti_dir, last));
        }
    }
    TRY("sync", mcg_synchronize(s));
    TRY("radial files", mcg_write_radial(s, mean_dir));
    mcg_destroy(s);
    return 0;
}
