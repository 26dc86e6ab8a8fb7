// xflat_run -- command-line host for the MI355X evolution step.
//
//   xflat_run <config.txt> [snapshot.nc] [--precision fp64|mixed|fp32]
//             [--module ma|maa|sa] [--chunk N] [--species 2|4] [--autotune N] [--quiet]
//
// --module stands for the reference's compile-time module choice (-DMA,
// -DMAA, -DSA, include/global.h:5-19); by default MAA when Pbins > 1, else MA.
//
// The reference's driver (src/main.cpp:11-38: argv[1] = config file,
// optional argv[2] = a Snapshot file to resume from) with its evolution loop
// (src/Nmr.cpp:89-262) and output plugin (src/IO_f.cpp) over the C-ABI:
//  * the configuration is read by xf_parse_config (src/Parser.cpp:25-127);
//  * a resume file sets R0, dr and psi from its last record
//    (io_f::fillInitData, IO_f.cpp:197-250), else pure flavour states;
//  * the loop runs while r < Rn, the wall time is below Tn and fewer than Ts
//    steps were attempted (Nmr.cpp:130-134), in launches of up to N steps
//    that stop on the device after an accepted step at which a dump is due;
//  * dumps follow io_f::dumpToFile (IO_f.cpp:387-424): Snapshot records
//    (dumpMode bit 0) and Probability records (bit 1) when r == Rn, r moved
//    more than r_step since the last record, it % itr_step == 0 (it = the
//    attempt index), or more than t_step seconds passed.  The wall-clock
//    trigger is evaluated between launches; new files every newFilestep
//    Snapshot records (IO_f.cpp:269-273); without IO the radius is printed
//    every 1 km (IO_f.cpp:391-398).
// One process drives one GPU (the reference's MPI rank 0 of 1).
#include <sys/time.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "xflat.h"
#include "xflat_io.h"

namespace {

double now_s() {
    timeval t;
    gettimeofday(&t, nullptr);
    return t.tv_sec + 1e-6 * t.tv_usec;
}

[[noreturn]] void die(const char* what, int rc) {
    std::fprintf(stderr, "xflat_run: %s failed (%d): %s\n", what, rc, xf_last_error());
    std::exit(-10);          // the reference's exit code on IO failure (IO_f.cpp:88)
}

#define CHECK(call)                         \
    do {                                    \
        int rc_ = (call);                   \
        if (rc_ != XF_OK) die(#call, rc_);  \
    } while (0)

// One output stream of io_f: its thresholds and state.
struct Stream {
    bool on = false;
    double rstep = 0.;       // r_step
    int tstep = 0;           // t_step [s] (0: none)
    int itrstep = 0;         // itr_step (0: none)
    double last_r = 0.;
    double start_t = 0.;     // whole seconds, as the reference's timeval tv_sec
    bool due(double r, double Rn, long long it, double t) const {
        if (!on) return false;
        if (r == Rn || r - last_r > rstep) return true;
        if (tstep > 0 && (long long)t - (long long)start_t > tstep) return true;
        if (itrstep > 0 && it % itrstep == 0) return true;
        return false;
    }
};

}  // namespace

int main(int argc, char** argv) {
    const char* cfg = nullptr;
    const char* init = nullptr;
    int precision = XF_FP64, species = 0, chunk = 2000, module = -1, autotune = 0;
    bool quiet = false;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        if (a == "--precision" && i + 1 < argc) {
            std::string v = argv[++i];
            precision = v == "mixed" ? XF_MIXED : v == "fp32" ? XF_FP32 : XF_FP64;
        } else if (a == "--module" && i + 1 < argc) {
            std::string v = argv[++i];
            module = v == "sa" ? XF_SA : v == "maa" ? XF_MAA : v == "ma" ? XF_MA : -2;
            if (module == -2) {
                std::fprintf(stderr, "xflat_run: --module must be ma, maa or sa\n");
                return 2;
            }
        } else if (a == "--chunk" && i + 1 < argc) {
            chunk = std::max(1, std::atoi(argv[++i]));
        } else if (a == "--species" && i + 1 < argc) {
            species = std::atoi(argv[++i]);
        } else if (a == "--autotune" && i + 1 < argc) {
            autotune = std::max(0, std::atoi(argv[++i]));
        } else if (a == "--quiet") {
            quiet = true;
        } else if (!cfg) {
            cfg = argv[i];
        } else if (!init) {
            init = argv[i];
        } else {
            std::fprintf(stderr, "xflat_run: unexpected argument %s\n", argv[i]);
            return 2;
        }
    }
    if (!cfg) {
        std::fprintf(stderr, "usage: xflat_run <config.txt> [snapshot.nc] [--precision fp64|mixed|fp32] "
                             "[--module ma|maa|sa] [--chunk N] [--species 2|4] [--autotune N] [--quiet]\n");
        return 2;
    }

    xf_params p;
    CHECK(xf_parse_config(cfg, &p));
    if (module >= 0) p.module = module;
    xf_opts o;
    std::memset(&o, 0, sizeof o);
    o.precision = precision;
    o.species = species;
    o.renorm = -1;
    o.resident = -1;
    xf_ctx* ctx = nullptr;
    CHECK(xf_create(&p, &o, &ctx));

    const size_t ncell = (size_t)xf_beam_len(ctx) * 4 * p.ebins;    // beams x prtcl x comp x ebin
    std::vector<double> psi(ncell);
    double r = p.R0, dr = p.dr;
    if (init) {
        std::printf("Provided data file: %s\n", init);
        std::printf("Loading the provided data file...\n");
        CHECK(xf_snapshot_read_last(init, &p, psi.data(), &r, &dr));
        p.R0 = r;
        p.dr = dr;
        CHECK(xf_state_upload(ctx, psi.data(), r, dr));
    } else {
        CHECK(xf_state_init_pure(ctx));
    }
    if (autotune > 0) {
        // time every launch shape over `autotune` steps from this state (which
        // the autotuner restores) and keep the fastest
        xf_opts best;
        double ms = 0.;
        CHECK(xf_autotune(ctx, autotune, &best, &ms));
        if (!quiet)
            std::printf("Launch shape: mode %d, %d cells per thread, %d threads (%.4f ms per step)\n", best.mode,
                        best.cpt, best.block, ms);
    }

    std::vector<double> fv(4 * (size_t)p.ebins), hvac(2 * (size_t)p.ebins), le(4), ev(4);
    CHECK(xf_get_spectra(ctx, fv.data(), hvac.data(), le.data(), ev.data()));

    // ---- outputs (io_f::init, IO_f.cpp:90-178) -----------------------------
    const double t0 = now_s();
    Stream s1, s2, s0;
    s1.on = (p.dump_mode & 1) != 0;
    s1.rstep = p.r_step1; s1.tstep = p.t_step1; s1.itrstep = p.itr_step1;
    s2.on = (p.dump_mode & 2) != 0;
    s2.rstep = p.r_step2; s2.tstep = p.t_step2; s2.itrstep = p.itr_step2;
    s0.on = p.dump_mode == 0;
    s0.rstep = 1.;                                  // rstep0, IO_f.cpp:38
    for (Stream* s : {&s1, &s2, &s0}) { s->last_r = p.R0; s->start_t = (double)(long long)t0; }
    int ctr = 0;
    long long snap_itr = 0;
    xf_nc* snap = nullptr;
    xf_nc* prob = nullptr;
    char name[1024];
    auto open_snapshot = [&]() {
        CHECK(xf_output_name(&p, 0, ctr, 0, name, sizeof name));
        CHECK(xf_snapshot_create(&p, name, &snap));
        snap_itr = 0;
    };
    if (s1.on) open_snapshot();
    if (s2.on) {
        CHECK(xf_output_name(&p, 1, ctr, 0, name, sizeof name));
        CHECK(xf_probability_create(&p, name, &prob));
    }

    // ---- the radial loop (Nmr.cpp:125-252) ---------------------------------
    long long total = 0, accepted = 0;
    double t_last = now_s();
    while (true) {
        t_last = now_s();
        const bool go = (r < p.Rn) && ((long long)t_last - (long long)t0 < p.Tn) && (total < p.Ts);
        if (!go) break;
        // device-side stop: the earliest radius trigger and every iteration
        // trigger (gcd of the active itr_steps; the host re-checks each)
        xf_stop st;
        std::memset(&st, 0, sizeof st);
        double next_r = 1e300;
        for (Stream* s : {&s1, &s2, &s0}) {
            if (!s->on) continue;
            if (s->last_r + s->rstep < next_r) { next_r = s->last_r + s->rstep; st.dump_last_r = s->last_r; st.dump_rstep = s->rstep; }
            if (s != &s0 && s->itrstep > 0) st.dump_itrstep = st.dump_itrstep ? std::gcd(st.dump_itrstep, s->itrstep) : s->itrstep;
        }
        if (st.dump_rstep <= 0. && (s1.on || s2.on || s0.on)) { st.dump_rstep = 1e-300; }
        const long long budget = std::min<long long>(chunk, p.Ts - total);
        xf_run_info info;
        CHECK(xf_run_steps(ctx, (int)budget, &st, nullptr, &info));
        total += info.steps;
        accepted += info.accepted;
        r = info.r;
        dr = info.dr;
        const double t = now_s();
        const long long it = total - 1;             // attempt index of the last attempt
        bool dumped_state = false;
        auto state = [&]() {
            if (!dumped_state) {
                double rr, dd;
                CHECK(xf_state_download(ctx, psi.data(), &rr, &dd));
                dumped_state = true;
            }
        };
        // Dumps happen after accepted steps only.  After a device stop for a
        // dump (stop_reason 2) the last attempt was accepted and each stream
        // re-checks all its triggers there; otherwise only the wall-clock
        // triggers can be due, for the state after the last accepted step.
        const bool stopped = info.stop_reason == 2;
        auto time_due = [&](const Stream& s) {
            return s.on && s.tstep > 0 && (long long)t - (long long)s.start_t > s.tstep;
        };
        auto due = [&](const Stream& s) {
            return stopped ? s.due(r, p.Rn, it, t) : (info.accepted > 0 && time_due(s));
        };
        if (due(s1)) {
            if (p.new_file_step > 0 && snap_itr >= p.new_file_step) {      // IO_f.cpp:269-273
                CHECK(xf_nc_close(snap));
                ++ctr;
                open_snapshot();
            }
            state();
            if (!quiet) std::printf("%f\n", r);
            CHECK(xf_snapshot_append(snap, r, info.last_dr, psi.data()));
            ++snap_itr;
            if (p.sync_step > 0 && snap_itr % p.sync_step == 0) CHECK(xf_nc_sync(snap));
            s1.last_r = r;
            s1.start_t = (double)(long long)t;
        }
        if (due(s2)) {
            state();
            if (!quiet) std::printf("%f\n", r);
            CHECK(xf_probability_append(prob, r, psi.data(), fv.data()));
            s2.last_r = r;
            s2.start_t = (double)(long long)t;
        }
        if (stopped && s0.on && r - s0.last_r > s0.rstep) {
            s0.last_r = r;
            if (!quiet) std::printf("%f\n", r);
        }
        if (info.stop_reason == 3) die("xf_run_steps (device error)", XF_EBARRIER);
        if (info.steps == 0) break;
    }
    std::printf("Total iterations: %lld\tCorrect iterations: %lld\t at radius(Km): %f\n", total, accepted, r);
    std::printf("total time:%f\n", now_s() - t0);
    if (snap) CHECK(xf_nc_close(snap));
    if (prob) CHECK(xf_nc_close(prob));
    CHECK(xf_destroy(ctx));
    return 0;
}
