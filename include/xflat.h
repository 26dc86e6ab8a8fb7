/*
 * xflat.h -- C-ABI of the MI355X multi-angle evolution step.
 *
 * The reference exposes this path as the compile-time plugin namespace
 * nbgrp::PHY_MODULE (include/Phy.h:18-161, selected by global.h:5-19) over the
 * NBeam class (include/NBeam.h:57-189), driven by nmr::evolutionLoop
 * (src/Nmr.cpp:89-262).  This library replaces it with plain C entry points:
 * opaque context handles that own device-resident structure-of-arrays beam
 * storage, int status returns (no exceptions cross the ABI) and plain
 * pointers/sizes only.
 *
 * Host arrays of wave functions always use the reference's NetCDF snapshot
 * layout psi[beam][prtcl][comp][ebin] (src/IO_f.cpp:276-292), where
 * beam = theta*Pbins + phi (MA: Pbins = 1), prtcl in {nu_e, anti-nu_e, nu_x,
 * anti-nu_x} (NBeam.h:19-22 and nbm_prtcls = 2*Flvs, NBeam.cpp:149-152) and
 * comp in {Re a, Im a, Re b, Im b} (NBeam::psi, NBeam.h:108-124).
 */
#ifndef XFLAT_H_
#define XFLAT_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define XF_OK          0
#define XF_EINVAL     -1   /* bad argument or unsupported configuration        */
#define XF_EHIP       -2   /* a HIP runtime call failed (message: xf_last_error) */
#define XF_ENOMEM     -3   /* device or host allocation failed                  */
#define XF_ESTATE     -4   /* state not SU(2)-partner symmetric for 2-species   */
#define XF_EIO        -5   /* file could not be read or written                 */
#define XF_EBARRIER   -6   /* a grid barrier did not complete (bounded wait)    */
#define XF_ENODEV     -7   /* no HIP device available                           */

/* ---- geometry modules (reference: PHY_MODULE choice, global.h:5-19) ---- */
#define XF_MA   0          /* bulb multi-angle, src/Phy_MA.cpp                  */
#define XF_MAA  1          /* multi-azimuthal-angle, src/Phy_MAA.cpp            */
#define XF_SA   2          /* single-angle bulb, src/Phy_SA.cpp: one trajectory,
                              H_vv = coef * D(r) * X, dl = dr (Phy_SA.cpp:144-147,250-283) */

/* ---- precision of the fused stepper ------------------------------------ */
#define XF_FP64   0        /* FP64 state and arithmetic (parity mode)           */
#define XF_MIXED  1        /* FP64 state; FP32 propagator increments;
                              FP64 moments, error norm, accept/reject          */
#define XF_FP32   2        /* FP32 state and propagator; FP64 moments, error    */

/*
 * Physics and run parameters, in the reference's internal units (after the
 * unit scaling of src/Parser.cpp:74-123).  xf_parse_config fills it from a
 * config*.txt file; unset keys are zero, as for the reference's
 * zero-initialised Util globals (src/Util.cpp:10-78).
 */
typedef struct xf_params {
    int    module;                 /* XF_MA, XF_MAA or XF_SA (SA: one beam,
                                      Abins= and Pbins= not used for the state) */
    int    abins, pbins, ebins;    /* Abins= Pbins= Ebins=                     */
    int    flvs;                   /* Flvs= (only 2 is supported, as upstream) */
    int    spoints;                /* SPoints= (parsed, unused by MA/MAA)      */
    double E0, E1;                 /* E0= E1= [1/km]  (MeV * MEV)              */
    double R0, Rn, dr, max_dr;     /* R0= Rn= dr= max_dr= [km]                 */
    double eps0, ksi, mu;          /* eps0= ksi= mu=                           */
    int    Ts, Tn;                 /* Ts= step budget, Tn= wall-time budget [s] */
    double theta, theta2;          /* theta= ; theta2 = 2*theta                 */
    double dm2;                    /* dm2= [1/km^2] (eV^2 * EV2)                */
    double L[4];                   /* Lve= Lv_e= Lvt= Lv_t= (erg/s * ERG_S)     */
    double T[4];                   /* Tve= Tv_e= Tvt= Tv_t= (MeV * MEV)         */
    double eta[4];                 /* eta_ve= eta_v_e= eta_vt= eta_v_t=         */
    int    has_matter;             /* hasMatter=                               */
    double Ye, Rv, Mns, gs, S, hNS;/* Ye= Rv= Mns= gs= S= hNS=                  */
    double nb0;                    /* nb0= [1/km^3] (cm^-3 * 1e15)              */
    /* IO and orchestration keywords (src/Parser.cpp:40-60) */
    int    dump_mode;              /* dumpMode= bit0 snapshot, bit1 probability */
    int    new_file_step, sync_step;
    double r_step1; int t_step1, itr_step1;
    double r_step2; int t_step2, itr_step2;
    int    start_beam, end_beam;
    int    multi_node_bench, min_nodes, max_nodes;
    char   file_prefix[256];
} xf_params;

/* Launch/precision options of the fused stepper.  species, block, cpt and
 * grid: 0 = automatic; renorm and resident: -1 = automatic (0 turns
 * renormalisation off, resp. forces the streaming modes). */
typedef struct xf_opts {
    int precision;        /* XF_FP64 / XF_MIXED / XF_FP32                       */
    int species;          /* 0 auto, 2 = SU(2)-partner reduced, 4 = explicit
                             (FP64 and XF_MIXED; XF_EINVAL with XF_FP32)        */
    int renorm;           /* unitarity renormalisation of accepted states
                             (-1 auto: on for XF_MIXED/XF_FP32, off for FP64)   */
    int block;            /* threads per workgroup (64..1024), 0 auto          */
    int cpt;              /* energy cells per thread per tile (1,2,4), 0 auto   */
    int resident;         /* 1 keep state in registers across steps, 0 stream,
                             -1 auto                                            */
    int grid;             /* workgroups (streaming mode), 0 auto                */
    int mode;             /* output only (xf_get_launch, xf_autotune): the
                             launch mode in use, 0 keep / 1 recompute (both
                             resident = 1), 2 / 4 stream (resident = 0)      */
} xf_opts;

/* One attempted radial step of the fused stepper (Nmr.cpp:153-249). */
typedef struct xf_step_rec {
    double  r;          /* radius at the start of the attempt                  */
    double  dr;         /* step length attempted                               */
    double  err;        /* sqrt(max over beams of sum_e fv |psi4-psi1|^2)      */
    int32_t accepted;   /* err < eps0                                          */
    int32_t pad;
} xf_step_rec;

/* Result of one xf_run_steps / xf_ensemble_run_steps call. */
typedef struct xf_run_info {
    int64_t steps;          /* attempted steps in this call                    */
    int64_t accepted;       /* accepted steps in this call                     */
    int64_t total_steps;    /* counter_tot since the state was set             */
    int64_t total_accepted; /* counter since the state was set                 */
    double  r, dr;          /* radius and next step length after the call      */
    double  last_err;       /* error of the last attempted step                */
    double  last_dr;        /* step length of the last accepted step (the dr
                               the reference writes with a dump, Nmr.cpp:238)  */
    double  kernel_ms;      /* device time of the call (HIP events)            */
    int32_t stop_reason;    /* 0 budget, 1 r reached Rn, 2 dump due, 3 error   */
    int32_t pad;
} xf_run_info;

/* Stop conditions for xf_run_steps beyond the step budget. */
typedef struct xf_stop {
    double dump_rstep;     /* >0: stop after an accepted step with
                              r - dump_last_r > dump_rstep or r == Rn
                              (IO trigger, src/IO_f.cpp:402)                   */
    double dump_last_r;
    int    dump_itrstep;   /* >0: stop after an accepted step whose attempt
                              index it satisfies it % dump_itrstep == 0        */
    int    pad;
} xf_stop;

typedef struct xf_ctx xf_ctx;
typedef struct xf_ensemble xf_ensemble;

/* ---- library ------------------------------------------------------------ */
const char* xf_last_error(void);                 /* thread-local message     */
int  xf_version(void);
int  xf_device_count(int* n);

/* Config file reader, keywords and unit scaling of Parser::readLines
 * (src/Parser.cpp:25-127).  Unknown keys are ignored, '#' starts a comment. */
int  xf_parse_config(const char* path, xf_params* out);
/* Same, from an in-memory text. */
int  xf_parse_config_text(const char* text, xf_params* out);

/* ---- context: replaces phy::init / nmr::init / nbm::init ----------------
 * xf_create runs the set-up of nbm::init (NBeam.cpp:138-209: spectra fv,
 * H_vac), phy::init (Phy_MA.cpp:233-284, Phy_MAA.cpp:247-304: angle grids,
 * hvv coefficient, L/<E>), matt::init (Matt.cpp:22-29) and allocates the four
 * beam sets of nmr::init (Nmr.cpp:48-70) on the device, initialised to pure
 * flavour states like NBeam::NBeam(int) (NBeam.cpp:236-277).               */
int  xf_create(const xf_params* p, const xf_opts* o, xf_ctx** out);
int  xf_destroy(xf_ctx* ctx);                    /* phy::freemem, Phy_MA.cpp:290-302 */

/* Dimension queries, Phy.h:24-67. */
int  xf_beam_len(const xf_ctx* ctx);             /* phy::beamLen               */
int  xf_get_dim(const xf_ctx* ctx);              /* phy::getDim (5 MA, 6 MAA, 4 SA) */
int  xf_count_dim(const xf_ctx* ctx, size_t* count);   /* phy::countDim        */

/* Host-side set-up values (for checks and output headers). */
int  xf_get_spectra(const xf_ctx* ctx, double* fv /*4*NE*/, double* hvac /*2*NE*/,
                    double* L_E /*4*/, double* Ev /*4*/);
int  xf_get_geometry(const xf_ctx* ctx, double* cos2R /*Abins*/, double* coef /*1*/);
/* matt::getHm (Matt.cpp:58-64). */
int  xf_get_hm(const xf_ctx* ctx, double r, double* hm);

/* ---- fine-grained Phy.h mirror on the four beam sets ---------------------
 * Beam set ids 0..3 stand for nubm_0, nubm_1, nubm_3, nubm_4 (Nmr.cpp:42-46);
 * hvv ids 0..3 for h0, h1, h3, h4 (Nmr.cpp:105-109).  All FP64, 4 species.  */
int  xf_beams_upload(xf_ctx* ctx, int set, const double* psi);
int  xf_beams_download(xf_ctx* ctx, int set, double* psi);
int  xf_beams_init_pure(xf_ctx* ctx, int set);                  /* NBeam(int)        */
int  xf_calc_esum(xf_ctx* ctx, int set);                        /* NBeam::calcESum, NBeam.cpp:812-840 */
int  xf_get_esum(xf_ctx* ctx, int set, double* out /*beamLen*4*/);  /* NBeam::getESum */
int  xf_calc_angle_bins(xf_ctx* ctx, double r, int pt);         /* Phy.h:108        */
int  xf_calc_delta_ls(xf_ctx* ctx, double dr, int cur, int s, int e); /* Phy.h:117 */
int  xf_calc_hvv(xf_ctx* ctx, int set, int pt, int hvv);        /* Phy.h:125        */
int  xf_get_hvv(xf_ctx* ctx, int hvv, double* out /*beams*4*/);
int  xf_evolve(xf_ctx* ctx, int in, int pt, int hvv, double hm, int out);      /* Phy.h:147 */
int  xf_evolve_hvv(xf_ctx* ctx, int in, int pt, int hvv, double hm, int out);  /* Phy.h:148 */
int  xf_evolve_avg(xf_ctx* ctx, int in, int pt, int hvv, double hm, int out, int avg); /* Phy.h:149 */
int  xf_evolve_avg_inplace(xf_ctx* ctx, int pt, int hvv, double hm, int io, int avg); /* Phy.h:152 */
int  xf_avg_beam(xf_ctx* ctx, int in, int out);                 /* Phy.h:159        */
/* max over beams of NBeam::calcErrHvv(a vs b) (NBeam.cpp:903-938, loop
 * Nmr.cpp:195-206), returned before the square root; re-caches ESum(a).     */
int  xf_calc_err_hvv(xf_ctx* ctx, int a, int b, double* max_err);

/* ---- fused stepper: replaces nmr::evolutionLoop (Nmr.cpp:89-262) --------
 * The state lives on the device in the precision of xf_opts.precision.
 * xf_state_upload takes a full 4-species psi (reference layout) plus the
 * radius and step length; with 2-species mode the nu_x / anti-nu_x columns
 * must be SU(2) partners (-conj(b), conj(a)) of nu_e / anti-nu_e, else
 * XF_ESTATE (with species 0 in FP64 or XF_MIXED the context switches to 4
 * species and takes the state instead).  xf_state_download reconstructs all 4
 * species.                                                                */
int  xf_state_upload(xf_ctx* ctx, const double* psi, double r, double dr);
int  xf_state_init_pure(xf_ctx* ctx);            /* pure flavour, r = R0, dr = dr= */
int  xf_state_download(xf_ctx* ctx, double* psi, double* r, double* dr);
int  xf_state_counters(const xf_ctx* ctx, int64_t* total, int64_t* accepted);
/* Moves the end of the run, Rn= (Util::setRn, the loop condition r < Rn of
 * Nmr.cpp:130; the step that would pass it is clamped to end on it,
 * Nmr.cpp:249), so that a run can stop exactly on any radius and go on from
 * there.  A run that ended on the old Rn has dr = 0; it restarts with the
 * length of its last accepted step, the dr the reference stores with a dump
 * and resumes from (Nmr.cpp:238, IO_f.cpp:197-250).                        */
int  xf_set_rn(xf_ctx* ctx, double rn);
int  xf_state_species(const xf_ctx* ctx);        /* 2 or 4 in use           */
/* Attempt up to max_steps radial steps (fewer if r reaches Rn or a stop
 * condition fires).  log (optional) receives one record per attempt.       */
int  xf_run_steps(xf_ctx* ctx, int max_steps, const xf_stop* stop,
                  xf_step_rec* log, xf_run_info* info);
/* xf_run_steps without any device allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  No counterpart in the reference.  The same
 * kernels, steps and results as xf_run_steps on the same context (bit for
 * bit); the launch descriptor, the exchange rows and the step log live in
 * one static device array of XF_NOALLOC_SCRATCH_BYTES per GPU, held until
 * the process ends, instead of the context's own buffers.  The context's
 * state buffers are the inputs and outputs, as for xf_run_steps.
 * Preconditions: a launch shape chosen at set-up (xf_get_launch,
 * xf_set_launch or xf_autotune, which may allocate), at most
 * XF_NOALLOC_MAX_WORKGROUPS workgroups, and with a log at most
 * XF_NOALLOC_MAX_LOG attempts; otherwise XF_EINVAL.  Every call on a GPU
 * shares that GPU's array: calls that overlap in time on one GPU (two host
 * threads, or two streams) race and give undefined results.  Each call
 * clears what must start at zero (status, rows) before its stepper runs,
 * and forms psi0's moments there when the context holds none for its
 * launch shape. */
#define XF_NOALLOC_MAX_WORKGROUPS 1024
#define XF_NOALLOC_MAX_LOG        4096
#define XF_NOALLOC_SCRATCH_BYTES  4379648   /* 1 KiB descriptor + 4 KiB map + 128 KiB log + 4 MiB rows
                                               + 48 KiB moments' partial sums */
int  xf_run_steps_noalloc(xf_ctx* ctx, int max_steps, const xf_stop* stop,
                          xf_step_rec* log, xf_run_info* info);
/* The radial loop as a HIP graph: `launches` launches of up to max_steps
 * attempted steps each, replayed from one instantiated graph (built on first
 * use, kept while launches, max_steps and the launch shape stay).  The same
 * steps and results as `launches` calls of xf_run_steps without stop
 * conditions or log; info counts all of them.  No counterpart in the
 * reference (the reference's loop is host code, Nmr.cpp:125-252). */
int  xf_run_graph(xf_ctx* ctx, int launches, int max_steps, xf_run_info* info);
/* Launch shape in use (after automatic choice). */
int  xf_get_launch(const xf_ctx* ctx, xf_opts* out, int* tiles, int* beams_per_tile);
/* Explicit launch shape: mode 0 keeps psi0 and the step's intermediate
 * states in registers for the whole launch, 1 keeps psi0 only and recomputes,
 * 2 streams the state from HBM, 4 streams it with the beam records of a
 * workgroup's tiles kept in LDS (for the whole step when they fit, else in
 * chunks); cpt energy cells per thread; block threads; bpt beams per tile; a
 * streaming grid is capped by xf_opts.grid.  No counterpart in the reference
 * (a GPU launch knob). */
int  xf_set_launch(xf_ctx* ctx, int mode, int cpt, int block, int bpt);
/* Registers, scratch bytes per lane, co-resident workgroups and LDS bytes of
 * the launch shape in use. */
int  xf_launch_info(const xf_ctx* ctx, int* vgprs, int* scratch, int* max_blocks, int* lds_bytes);
/* Profiling: per-phase real-time stamps (100 MHz counter) of workgroup 0
 * for the first `steps` steps of the next xf_run_steps (16 slots per step:
 * [0] step start; [1..4] pass A: cells done, partial published, arrived,
 * totals known; [5..8] pass B; [14] pass C's cells done, [15] its error
 * sums done, [9..12] its exchange; [13] decision).  Tracing changes no
 * result.  Arm with
 * xf_trace_next(ctx, buf, steps), run, then xf_trace_fetch(ctx, buf, steps).
 * No counterpart in the reference. */
int  xf_trace_next(xf_ctx* ctx, long long* out, int steps);
int  xf_trace_fetch(xf_ctx* ctx, long long* out, int steps);
/* Project autotuner: times candidate (block, cpt, resident) shapes for
 * `steps` steps each on a scratch copy of the state and keeps the fastest. */
int  xf_autotune(xf_ctx* ctx, int steps, xf_opts* best, double* best_ms_per_step);

/* ---- batched ensemble: many independent instances, one persistent launch -
 * Every instance has its own grid, spectra, matter profile, r, dr and
 * accept/reject; workgroups are partitioned between instances.            */
int  xf_ensemble_create(const xf_params* p, int n, const xf_opts* o, xf_ensemble** out);
int  xf_ensemble_destroy(xf_ensemble* ens);
int  xf_ensemble_upload(xf_ensemble* ens, int i, const double* psi, double r, double dr);
int  xf_ensemble_init_pure(xf_ensemble* ens, int i);
int  xf_ensemble_download(xf_ensemble* ens, int i, double* psi, double* r, double* dr);
int  xf_ensemble_run_steps(xf_ensemble* ens, int max_steps, xf_run_info* info /*n*/);
/* xf_run_graph for an ensemble (one graph launches all instances). */
int  xf_ensemble_run_graph(xf_ensemble* ens, int launches, int max_steps, xf_run_info* info /*n*/);
/* Launch-shape tuner over all instances (as xf_autotune; states restored). */
int  xf_ensemble_autotune(xf_ensemble* ens, int steps, xf_opts* best, double* best_ms_per_step);
int  xf_ensemble_cells(const xf_ensemble* ens, int64_t* cells);
/* xf_set_rn for instance i of an ensemble (each instance stops on its own Rn). */
int  xf_ensemble_set_rn(xf_ensemble* ens, int i, double rn);
int  xf_ensemble_launch(const xf_ensemble* ens, int* mode, int* cpt, int* block, int* grid);

#ifdef __cplusplus
}
#endif

#endif /* XFLAT_H_ */
