// Runs the dam break on the device through the C-ABI alone, as the
// reference's main program runs it (src/main.cpp:33-135): the initial
// condition, then pressure, force, Shepard filter and gravity
// (wcsph_sim_init), then steps to end_time with a snapshot and a restart
// file each time the next output time is reached (OutputFileWithTimeInterval,
// src/io.h:2-37). Given a restart file it resumes from that state instead
// (src/main.cpp:45-47), continuing the output count at k0.
//
// usage: sim_driver dim dx end_time n_snapshots outdir precision shepard_interval [restart.bin k0]
//   precision: 0 FP32, 1 mixed, 2 FP64
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "wcsph_abi.h"

static int fail(const char* what) {
    fprintf(stderr, "sim_driver: %s: %s\n", what, wcsph_last_error());
    return 3;
}

static std::string name(const std::string& dir, int k, const char* ext) {
    char buf[64];
    snprintf(buf, sizeof buf, "/%05d_%05d_%05d.%s", k, 1, 0, ext);  // FDPS per-rank naming, one process
    return dir + buf;
}

int main(int argc, char** argv) {
    if (argc != 8 && argc != 10) {
        fprintf(stderr, "usage: sim_driver dim dx end_time n_snapshots outdir precision shepard_interval "
                        "[restart.bin k0]\n");
        return 2;
    }
    const int dim = atoi(argv[1]);
    const double dx = atof(argv[2]), end_time = atof(argv[3]);
    const int nsnap = atoi(argv[4]);
    const std::string outdir = argv[5];
    const int precision = atoi(argv[6]), shepard = atoi(argv[7]);
    const char* restart = argc == 10 ? argv[8] : nullptr;
    const int k0 = argc == 10 ? atoi(argv[9]) : 0;

    wcsph_sim* s = nullptr;
    if (restart) {
        if (wcsph_sim_restore(restart, 0, &s) != WCSPH_OK) return fail("restore");
    } else {
        // physics of the dam-break configurations
        wcsph_params p;
        if (wcsph_params_reference(dim, &p) != WCSPH_OK) return fail("params");
        p.alpha = 0.1;
        p.beta = 0.0;
        p.c_cfl = 0.25;
        p.g = 9.81;
        p.rho0 = 1000.0;
        p.gamma = 7.0;
        p.nu = 1.0e-6;
        p.c0 = 10.0 * std::sqrt(2.0 * p.g * 2.0);
        p.eps = 0.01;
        p.vsig_mode = 0;
        p.rho_floor = 0.0;
        p.shepard_interval = shepard;
        wcsph_dam_break_desc d = {dim, 3, dx, 1.0, 2.0, 0.5, 4.0, 3.0, 1.3};
        int64_t nf = 0, nw = 0;
        if (wcsph_dam_break_count(&d, &nf, &nw) != WCSPH_OK) return fail("count");
        const int64_t n = nf + nw;
        std::vector<double> pos(dim * n), vel(dim * n), dens(n), mass(n), h(n);
        std::vector<int32_t> type(n);
        if (wcsph_dam_break_fill(&d, &p, pos.data(), vel.data(), dens.data(), mass.data(), h.data(),
                                 type.data()) != WCSPH_OK)
            return fail("fill");
        wcsph_sim_config cfg = {};
        cfg.precision = precision;
        if (wcsph_sim_create(&p, &cfg, n, &s) != WCSPH_OK) return fail("create");
        if (wcsph_sim_upload(s, pos.data(), vel.data(), dens.data(), mass.data(), h.data(), type.data()) !=
            WCSPH_OK)
            return fail("upload");
        if (wcsph_sim_init(s) != WCSPH_OK) return fail("init");
    }

    // output times accumulate as in the reference: t_next += end_time / N
    const double interval = end_time / nsnap;
    double t_next = 0.0;
    for (int k = 0; k < k0; ++k) t_next += interval;
    int k = k0;
    wcsph_sim_stats st;
    for (;;) {
        if (wcsph_sim_get_stats(s, &st) != WCSPH_OK) return fail("stats");
        if (st.time >= t_next) {
            if (wcsph_sim_write_ascii(s, name(outdir, k, "dat").c_str()) != WCSPH_OK) return fail("snapshot");
            if (wcsph_sim_save(s, name(outdir, k % 10, "bin").c_str()) != WCSPH_OK) return fail("save");
            t_next += interval;
            ++k;
        }
        if (st.time >= end_time) break;
        // queue about half the steps left to the next output time, then single steps
        const double left = std::fmin(t_next, end_time) - st.time;
        const int chunk = st.dt > 0.0 ? (int)(0.5 * left / st.dt) : 0;
        if (wcsph_sim_step(s, chunk > 1 ? chunk : 1) != WCSPH_OK) return fail("step");
    }
    printf("snapshots %d..%d time %.9e step %lld pairs %lld\n", k0, k - 1, st.time, (long long)st.step,
           (long long)st.pairs_total);
    wcsph_sim_destroy(s);
    return 0;
}
