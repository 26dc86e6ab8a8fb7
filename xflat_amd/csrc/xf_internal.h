// xf_internal.h -- shared by the host set-up (xf_setup.cpp), the Phy mirror
// kernels (xf_phy.hip) and the fused stepper (xf_step.hip).
#pragma once

#include <cstdint>
#include <hip/hip_runtime_api.h>
#include <string>
#include <vector>

#include "../../include/xflat.h"


// Constants of include/Util.h:15-22 and include/global.h:43-46.  The
// reference truncates pi and sqrt(2); FP64 parity needs the same values.
namespace xfc {
constexpr double MEV = 5.067731161976211e+15;
constexpr double MEV_2 = 3.893793036626876e-32;
constexpr double ERG_S = 1.053613095584080e+16;
constexpr double EV2 = 2.568189913006477e+19;
constexpr double SQRT2 = 1.41421356237;
constexpr double PI = 3.14159265359;
constexpr double GF = 1.16637E-11 * MEV_2;
constexpr int NPR = 4;      // particles nu_e, anti-nu_e, nu_x, anti-nu_x
}  // namespace xfc

// Error reporting (thread-local last message).
int xf_fail(int code, const char* fmt, ...);

// ---- host-side set-up tables (nbm::init, phy::init, matt::init) ----------
struct XfTables {
    int module, nth, nph, ne, nbeam;
    double dE;
    std::vector<double> fv;        // [4][ne] normalised spectra (NBeam.cpp:154-208)
    std::vector<double> hvac;      // [ne][2] (NBeam.cpp:180-184)
    double Ev[4], L_E[4];
    std::vector<double> cos2R, dcos2R, sinR;   // [nth] (Phy_MA.cpp:141-164, Phy_MAA.cpp:160-174)
    std::vector<double> cosphi, sinphi;        // [nph] (Phy_MAA.cpp:177-185)
    double dPhi;
    double integ_coef;             // Phy_MA.cpp:271-272 / Phy_MAA.cpp:297-298
    double mat_coeff, nbfar_coeff; // Matt.cpp:22-29
};
int xf_build_tables(const xf_params* p, XfTables* t);
double xf_host_hm(const xf_params* p, const XfTables* t, double r);

// ---- device descriptor of one instance for the fused stepper --------------
// Moment layout: per source point NM = 2 (MA: J0, J1) or 4 (MAA: J0, Jx, Jy,
// Jz) moments times 3 density components (0, 2, 3; component 1 is zero).
constexpr int XF_M0MAX = 16;                   // moments carried between launches (NM * 3 <= 12)
struct XfRunScalars {
    double r, dr;
    long long total, accepted;     // counter_tot, counter (Nmr.cpp:123)
    double last_err;
    double last_dr;                // dr of the last accepted step
    int cur;                       // state buffer holding psi0 (streaming)
    int status;                    // 0 ok, XF_EBARRIER ...
    long long steps_call, acc_call;
    int stop_reason;
    unsigned gen0;                 // first exchange generation of the next launch
    // psi0's angular moments at radius m0_r, left by the last launch for the
    // next one of the same shape (m0_key): the cached energy sums that
    // Nmr.cpp:117-121 carries from step to step, so a launch resumes where
    // the previous one stopped instead of recomputing them
    int m0_ok, pad;
    long long m0_key;
    double m0_r;
    double m0[XF_M0MAX];
    // the first exchange wait that expired: where (1 poll, 2 fetch; 0 none),
    // the workgroup, the tag it waited for, the rows missing and the tag the
    // first missing row held (a row that never came holds an older tag)
    int diag_where, diag_wg;
    unsigned diag_tag, diag_seen;
    int diag_missing, diag_row;
};

// Cross-workgroup reduction of one instance (all-gather of tagged rows):
// each workgroup publishes its partial sums as one row of 64-bit words, each
// word carrying the exchange's generation + 1: the FP64 stepper splits value
// v into its low and high 32-bit halves (words 2v, 2v + 1) under a 32-bit
// tag, the FP32-increment steppers put value v in word v with the tag's three
// low bits in its three low mantissa bits.  Rows are value-major,
// rows[parity][word][nblk], so word w of all rows is one contiguous array.  Every workgroup fetches every row and reduces them in
// the same order.  Generations continue from launch to launch (the run
// scalars' gen0): a slot then holds the tag of its current exchange or of the
// one two before it, never a later one.  The rows are cleared and the count
// restarted at 0 only where that chain breaks (new rows or workgroup map, a
// new state, a failed launch, the no-allocation path's own rows) or where the
// tags would pass XF_MAX_EXCHANGES.
constexpr int XF_NVMAX = 32;                   // max values per partial row
constexpr int XF_ROWW = 2 * XF_NVMAX;          // words per row (at most)
constexpr int XF_ROW_REPLICAS = 4;             // copies of every row (reads spread over more lines)
constexpr long long XF_MAX_EXCHANGES = 0xF0000000ll;

struct XfInstDev {
    int module, nth, nph, ne, nbeam;
    int ns;                        // 2 (partner-reduced) or 4 species stored
    int bpt, ntiles;               // beams per tile, tiles
    int blk0, nblk;                // workgroups [blk0, blk0+nblk) of the launch
    long long ncells;              // nbeam * ne
    // tables
    const double* x;               // cos^2 theta at R_nu [nth]
    const double* dx;              // its bin width [nth]
    const double* sinR;            // [nth]
    const double* cphi;            // [nph]
    const double* sphi;            // [nph]
    const double* hv;              // [2][ne] H_vac components (h_r0, h_r1)
    const double* srcw;            // [ns][ne] source weights (L/<E> fv combos)
    const double* errw;            // [4][ne] error weights (fv of the 4 particles)
    const float* hvf;              // float copies
    const float* srcwf;            // srcw / its max |.| (L/<E> ~ 1e50 exceeds float)
    const float* errwf;
    // scalars
    double Rv, coef, mat_coeff, nbfar_coeff, nb0, hNS;
    double coef_f;                 // coef * source-weight scale of the float tables
    int has_matter, renorm;
    double eps0, ksi, max_dr, Rn;
    // state and work buffers
    void* psi[2];                  // [ns*4][ncells] planes, double or float
    XfRunScalars* rs;
    XfRunScalars* rs_host;         // pinned host copy, written by workgroup 0 as a launch ends
    unsigned long long* rows;      // [2][XF_ROW_REPLICAS][XF_ROWW][nblk] tagged partial rows
    xf_step_rec* log;             // optional per-attempt log
    long long* trace;              // optional [steps][16] real-time stamps of workgroup 0
    int trace_steps;
    int pad2;
    // stop predicates
    double dump_rstep, dump_last_r;
    int dump_itrstep;
    int tchunk;                    // mode 4: tiles whose beam records LDS holds at once
    int pad3;
    // the launch shape (workgroups, block, tile, chunk, cells per lane,
    // mode), which fixes how the moments' partial sums associate: carried
    // moments are used by launches of the same shape only
    long long launch_key;
};

// ---- launch configuration ---------------------------------------------------
struct XfLaunch {
    int precision;     // XF_FP64 / XF_MIXED / XF_FP32
    int ns;            // 2 or 4
    int mode;          // 0 resident-keep, 1 resident-recompute, 2 streaming, 4 streaming with the
                       // workgroup's beam records kept in LDS for the step
    int cpt;           // 1, 2, 4
    int block;         // threads per workgroup
    int bpt;           // beams per tile
    int grid;          // workgroups (all instances)
    int wbeams;        // mode 4: largest beam count of one workgroup's tiles (records in LDS)
    int ne;            // largest energy-bin count of the launch's instances (per-energy LDS table)
};

constexpr size_t XF_LDS_MAX = 160 * 1024;       // LDS per workgroup (and per CU) on gfx950
constexpr size_t XF_LDS_GRANULE = 2048;         // LDS per workgroup counted in these units for co-residency

// psi0's moments of every instance of a launch into its run scalars, keyed to
// the instance's launch shape (xf_moments.hip): nparts slices per instance,
// partial sums in d_part [ninst][nparts][NM * 3].
constexpr int XF_MOM_PARTS = 512;
int xf_moments_launch(const XfLaunch& L, int geom, const XfInstDev* d_insts, int ninst, int nparts, double* d_part,
                      void* stream);

// Launch helpers of the fused stepper (xf_step_host.hip).
bool xf_step_has_shape(const XfLaunch& L, int geom);
size_t xf_step_smem_bytes(const XfLaunch& L, int geom);
int xf_step_max_records(const XfLaunch& L, int geom);
int xf_step_max_blocks_geom(const XfLaunch& L, int geom, int* max_blocks);
int xf_step_launch_geom(const XfLaunch& L, int geom, const XfInstDev* d_insts, const int* d_blk_inst, int max_steps,
                        void* stream);
int xf_step_func_info(const XfLaunch& L, int geom, int* vgprs, int* spill_bytes);
// Clears the run status and the exchange rows of every instance of a launch
// and restarts their generation count (a kernel, so the launch sequence holds
// no memset node).
int xf_step_prepare(const XfInstDev* d_insts, int ninst, void* stream);
// The no-allocation entry point's device memory (xf_noalloc.hip): the
// regions of the static scratch array on the current device.
struct XfNoAllocArea {
    XfInstDev* desc;               // launch descriptor of the one instance
    int* blk;                      // workgroup -> instance map (all 0)
    xf_step_rec* log;              // per-attempt log
    unsigned long long* rows;      // exchange rows for up to XF_NOALLOC_MAX_WORKGROUPS
    double* part;                  // moments' partial sums (XF_MOM_PARTS x 12)
};
int xf_noalloc_area(XfNoAllocArea* out);
// Writes descriptor `d` into the area and clears the map for `grid`
// workgroups (a kernel on `stream`: no copy or memset node).
int xf_noalloc_setup(const XfInstDev& d, const XfNoAllocArea& a, int grid, void* stream);
// The radial loop as an instantiated graph of `launches` launches of up to
// max_steps steps (the first without clearing: the caller clears before it).
int xf_step_build_graph(const XfLaunch& L, int geom, const XfInstDev* d_insts, const int* d_blk_inst, int ninst,
                        int max_steps, int launches, hipGraphExec_t* exec);
