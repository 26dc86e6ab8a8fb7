// xf_capi.cpp -- C-ABI of the fused stepper and the batched ensemble:
// device buffers, state upload/download, launch-shape choice, run control.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <queue>
#include <vector>

#include "xf_internal.h"
#include "xf_phy.h"

// launch helpers (xf_step_host.hip)
bool xf_step_has_shape(const XfLaunch& L, int geom);
size_t xf_step_smem_bytes(const XfLaunch& L, int geom);
int xf_step_max_blocks_geom(const XfLaunch& L, int geom, int* max_blocks);
int xf_step_launch_geom(const XfLaunch& L, int geom, const XfInstDev* d_insts, const int* d_blk_inst, int max_steps,
                        void* stream);
int xf_step_func_info(const XfLaunch& L, int geom, int* vgprs, int* spill_bytes);

#define HIPCHK(x)                                                                                   \
    do {                                                                                            \
        hipError_t e_ = (x);                                                                        \
        if (e_ != hipSuccess) return xf_fail(XF_EHIP, "%s: %s", #x, hipGetErrorString(e_));         \
    } while (0)

namespace {

template <class T> int dalloc(T** p, size_t n) {
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T));
    if (e != hipSuccess) return xf_fail(XF_ENOMEM, "hipMalloc(%zu B): %s", n * sizeof(T), hipGetErrorString(e));
    *p = (T*)q;
    return XF_OK;
}
template <class T> int dput(T** p, const std::vector<T>& v) {
    int rc = dalloc(p, v.size());
    if (rc) return rc;
    HIPCHK(hipMemcpy(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return XF_OK;
}
std::vector<float> to_f(const std::vector<double>& v) { return std::vector<float>(v.begin(), v.end()); }

}  // namespace

// One instance of the stepper: tables, state buffers and run scalars on the device.
struct XfInst {
    xf_params p{};
    XfTables t{};
    int ns = 2, precision = XF_FP64, renorm = 0;
    double src_scale = 1.;
    size_t elem = 8;
    double *d_x = nullptr, *d_dx = nullptr, *d_sinR = nullptr, *d_cphi = nullptr, *d_sphi = nullptr;
    double *d_hv = nullptr, *d_srcw = nullptr, *d_errw = nullptr;
    float *d_hvf = nullptr, *d_srcwf = nullptr, *d_errwf = nullptr;
    void* d_psi[2] = {nullptr, nullptr};
    XfRunScalars* d_rs = nullptr;
    // pinned host copy of the run scalars, written by the stepper's workgroup
    // 0 as a launch ends (h_rs on the host, hd_rs its device address): the
    // host reads a launch's outcome without a device-to-host copy
    XfRunScalars* h_rs = nullptr;
    XfRunScalars* hd_rs = nullptr;
    XfRunScalars last{};           // the copy as the last launch left it
    unsigned long long* d_rows = nullptr;
    int rows_nblk = 0;
    // the rows hold tags of this instance's generations below gen0 only (see
    // XF_MAX_EXCHANGES): the next launch may continue without clearing them
    bool rows_ok = false;
    unsigned gen0 = 0;
    // the launch shape for which the run scalars hold psi0's moments (0: none)
    long long m0_key = 0;
    XfInstDev dev{};

    ~XfInst() {
        void* ps[] = {d_x, d_dx, d_sinR, d_cphi, d_sphi, d_hv, d_srcw, d_errw, d_hvf, d_srcwf, d_errwf,
                      d_psi[0], d_psi[1], d_rs, d_rows};
        for (void* q : ps)
            if (q) (void)hipFree(q);
        if (h_rs) (void)hipHostFree(h_rs);
    }

    long long ncells() const { return (long long)t.nbeam * t.ne; }

    int init(const xf_params* pp, int prec, int nspec, int ren) {
        p = *pp;
        int rc = xf_build_tables(&p, &t);
        if (rc) return rc;
        precision = prec;
        ns = nspec;
        renorm = ren;
        elem = (precision == XF_FP32) ? 4 : 8;
        const int ne = t.ne;
        std::vector<double> hv(2 * ne), srcw((size_t)ns * ne), errw(4 * (size_t)ne);
        for (int e = 0; e < ne; ++e) {
            hv[e] = t.hvac[2 * e];
            hv[ne + e] = t.hvac[2 * e + 1];
            for (int s = 0; s < 4; ++s) errw[(size_t)s * ne + e] = t.fv[(size_t)s * ne + e];
            if (ns == 2) {
                srcw[e] = t.L_E[0] * t.fv[e] - t.L_E[2] * t.fv[2 * ne + e];
                srcw[ne + e] = t.L_E[1] * t.fv[ne + e] - t.L_E[3] * t.fv[3 * ne + e];
            } else {
                for (int s = 0; s < 4; ++s) srcw[(size_t)s * ne + e] = t.L_E[s] * t.fv[(size_t)s * ne + e];
            }
        }
        // float source weights are scaled to O(1); the scale moves into coef_f
        src_scale = 0.;
        for (double w : srcw) src_scale = std::max(src_scale, std::fabs(w));
        if (!(src_scale > 0.)) src_scale = 1.;
        std::vector<double> srcs(srcw.size());
        for (size_t i = 0; i < srcw.size(); ++i) srcs[i] = srcw[i] / src_scale;

        if ((rc = dput(&d_x, t.cos2R)) || (rc = dput(&d_dx, t.dcos2R)) || (rc = dput(&d_sinR, t.sinR)) ||
            (rc = dput(&d_cphi, t.cosphi)) || (rc = dput(&d_sphi, t.sinphi)) || (rc = dput(&d_hv, hv)) ||
            (rc = dput(&d_srcw, srcw)) || (rc = dput(&d_errw, errw)) || (rc = dput(&d_hvf, to_f(hv))) ||
            (rc = dput(&d_srcwf, to_f(srcs))) || (rc = dput(&d_errwf, to_f(errw))))
            return rc;
        size_t nst = (size_t)ns * 4 * ncells();
        for (int b = 0; b < 2; ++b) {
            void* q = nullptr;
            hipError_t e = hipMalloc(&q, nst * elem);
            if (e != hipSuccess) return xf_fail(XF_ENOMEM, "state buffer (%zu B): %s", nst * elem, hipGetErrorString(e));
            d_psi[b] = q;
        }
        if ((rc = dalloc(&d_rs, 1))) return rc;
        HIPCHK(hipMemset(d_rs, 0, sizeof(XfRunScalars)));
        {
            void* h = nullptr;
            hipError_t e = hipHostMalloc(&h, sizeof(XfRunScalars), hipHostMallocCoherent);
            if (e != hipSuccess) return xf_fail(XF_ENOMEM, "pinned run scalars: %s", hipGetErrorString(e));
            h_rs = (XfRunScalars*)h;
            std::memset(h_rs, 0, sizeof(XfRunScalars));
            void* hd = nullptr;
            e = hipHostGetDevicePointer(&hd, h, 0);
            if (e != hipSuccess) return xf_fail(XF_EHIP, "hipHostGetDevicePointer: %s", hipGetErrorString(e));
            hd_rs = (XfRunScalars*)hd;
        }
        fill_dev();
        return XF_OK;
    }

    // exchange rows for nblk workgroups: [2][XF_ROW_REPLICAS][XF_ROWW][nblk] words
    int ensure_rows(int nblk) {
        if (nblk <= rows_nblk) return XF_OK;
        if (d_rows) (void)hipFree(d_rows);
        d_rows = nullptr;
        rows_nblk = 0;
        int rc = dalloc(&d_rows, (size_t)2 * XF_ROW_REPLICAS * XF_ROWW * nblk);
        if (rc) return rc;
        rows_nblk = nblk;
        rows_ok = false;
        dev.rows = d_rows;
        return XF_OK;
    }

    void fill_dev() {
        XfInstDev& d = dev;
        d.module = t.module; d.nth = t.nth; d.nph = t.nph; d.ne = t.ne; d.nbeam = t.nbeam;
        d.ns = ns;
        d.ncells = ncells();
        d.x = d_x; d.dx = d_dx; d.sinR = d_sinR; d.cphi = d_cphi; d.sphi = d_sphi;
        d.hv = d_hv; d.srcw = d_srcw; d.errw = d_errw; d.hvf = d_hvf; d.srcwf = d_srcwf; d.errwf = d_errwf;
        d.Rv = p.Rv; d.coef = t.integ_coef; d.coef_f = t.integ_coef * src_scale; d.mat_coeff = t.mat_coeff; d.nbfar_coeff = t.nbfar_coeff;
        d.nb0 = p.nb0; d.hNS = p.hNS; d.has_matter = p.has_matter; d.renorm = renorm;
        d.eps0 = p.eps0; d.ksi = p.ksi; d.max_dr = p.max_dr; d.Rn = p.Rn;
        d.psi[0] = d_psi[0]; d.psi[1] = d_psi[1];
        d.rs = d_rs; d.rs_host = hd_rs; d.rows = d_rows;
        d.log = nullptr;
        d.trace = nullptr;
        d.trace_steps = 0;
        d.dump_rstep = 0; d.dump_last_r = 0; d.dump_itrstep = 0;
    }

    // psi[beam][prtcl][comp][e] (reference layout) -> device planes
    int upload(const double* psi, double r, double dr) {
        const long long nc = ncells();
        const int ne = t.ne;
        const size_t nst = (size_t)ns * 4 * nc;
        std::vector<double> pl(nst);
        double worst = 0;
        for (int b = 0; b < t.nbeam; ++b)
            for (int s = 0; s < ns; ++s)
                for (int c = 0; c < 4; ++c)
                    for (int e = 0; e < ne; ++e)
                        pl[((size_t)s * 4 + c) * nc + (size_t)b * ne + e] = psi[(((size_t)b * 4 + s) * 4 + c) * ne + e];
        if (ns == 2) {
            // nu_x = (-conj b, conj a) of nu_e (and likewise for the antineutrinos)
            for (int b = 0; b < t.nbeam; ++b)
                for (int s = 0; s < 2; ++s) {
                    const double* E = psi + (((size_t)b * 4 + s) * 4) * ne;
                    const double* X = psi + (((size_t)b * 4 + s + 2) * 4) * ne;
                    for (int e = 0; e < ne; ++e) {
                        worst = std::max(worst, std::fabs(X[0 * ne + e] + E[2 * ne + e]));
                        worst = std::max(worst, std::fabs(X[1 * ne + e] - E[3 * ne + e]));
                        worst = std::max(worst, std::fabs(X[2 * ne + e] - E[0 * ne + e]));
                        worst = std::max(worst, std::fabs(X[3 * ne + e] + E[1 * ne + e]));
                    }
                }
            if (!(worst <= 1e-12))
                return xf_fail(XF_ESTATE, "nu_x columns are not SU(2) partners of nu_e (max deviation %.3g)", worst);
        }
        if (elem == 8) {
            HIPCHK(hipMemcpy(d_psi[0], pl.data(), nst * 8, hipMemcpyHostToDevice));
        } else {
            std::vector<float> f(pl.begin(), pl.end());
            HIPCHK(hipMemcpy(d_psi[0], f.data(), nst * 4, hipMemcpyHostToDevice));
        }
        XfRunScalars rs{};
        rs.r = r; rs.dr = dr; rs.cur = 0;
        HIPCHK(hipMemcpy(d_rs, &rs, sizeof rs, hipMemcpyHostToDevice));
        rows_ok = false;                 // the generation count restarts with the state
        m0_key = 0;                      // and the moments are the new state's, not yet formed
        last = rs;
        return XF_OK;
    }

    int scalars(XfRunScalars* rs) const {
        HIPCHK(hipMemcpy(rs, d_rs, sizeof *rs, hipMemcpyDeviceToHost));
        return XF_OK;
    }

    int download(double* psi, double* r, double* dr) const {
        XfRunScalars rs;
        int rc = scalars(&rs);
        if (rc) return rc;
        const long long nc = ncells();
        const int ne = t.ne;
        const size_t nst = (size_t)ns * 4 * nc;
        std::vector<double> pl(nst);
        if (elem == 8) {
            HIPCHK(hipMemcpy(pl.data(), d_psi[rs.cur], nst * 8, hipMemcpyDeviceToHost));
        } else {
            std::vector<float> f(nst);
            HIPCHK(hipMemcpy(f.data(), d_psi[rs.cur], nst * 4, hipMemcpyDeviceToHost));
            std::copy(f.begin(), f.end(), pl.begin());
        }
        if (psi) {
            for (int b = 0; b < t.nbeam; ++b)
                for (int s = 0; s < ns; ++s)
                    for (int c = 0; c < 4; ++c)
                        for (int e = 0; e < ne; ++e)
                            psi[(((size_t)b * 4 + s) * 4 + c) * ne + e] = pl[((size_t)s * 4 + c) * nc + (size_t)b * ne + e];
            if (ns == 2)
                for (int b = 0; b < t.nbeam; ++b)
                    for (int s = 0; s < 2; ++s) {
                        const double* E = psi + (((size_t)b * 4 + s) * 4) * ne;
                        double* X = psi + (((size_t)b * 4 + s + 2) * 4) * ne;
                        for (int e = 0; e < ne; ++e) {
                            X[0 * ne + e] = -E[2 * ne + e];
                            X[1 * ne + e] = E[3 * ne + e];
                            X[2 * ne + e] = E[0 * ne + e];
                            X[3 * ne + e] = -E[1 * ne + e];
                        }
                    }
        }
        if (r) *r = rs.r;
        if (dr) *dr = rs.dr;
        return XF_OK;
    }
};

struct xf_ensemble {
    xf_opts o{};
    XfLaunch L{};
    int geom = XF_MA;
    std::vector<std::unique_ptr<XfInst>> inst;
    XfInstDev* d_insts = nullptr;
    int* d_blk_inst = nullptr;
    xf_step_rec* d_log = nullptr;
    int log_cap = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool shaped = false;
    std::vector<XfInstDev> shadow;      // what d_insts holds (refreshed only when it differs)
    double* d_part = nullptr;           // moments' partial sums [inst][XF_MOM_PARTS][NM * 3]
    // the radial loop as an instantiated graph (xf_run_graph), kept while its
    // key -- launches, steps per launch, launch shape, descriptor arrays -- holds
    hipGraphExec_t gexec = nullptr;
    int g_launches = 0, g_steps = 0;
    XfLaunch g_L{};
    const XfInstDev* g_insts = nullptr;
    const int* g_blk = nullptr;

    ~xf_ensemble() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (d_insts) (void)hipFree(d_insts);
        if (d_blk_inst) (void)hipFree(d_blk_inst);
        if (d_log) (void)hipFree(d_log);
        if (d_part) (void)hipFree(d_part);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

struct xf_ctx {
    xf_params p{};
    xf_opts o{};
    XfTables t{};
    xf_ensemble* ens = nullptr;       // the fused stepper (one instance)
    XfPhy* phy = nullptr;             // the Phy.h mirror (four FP64 beam sets)
};

// ---------------------------------------------------------------- shapes
// Scratch (register spills) a stepper launch, whose workgroups must all be
// resident at once, may use over its whole grid.
constexpr double XF_COOP_SCRATCH_MAX = 32.0 * 1048576.0;

static int max_bpt(int ne, int cpt) {
    int b = (1024 * cpt) / ne;
    return std::max(1, std::min(128, b));
}
static int block_for(int bpt, int ne, int cpt) {
    int cells = bpt * ne;
    int thr = (cells + cpt - 1) / cpt;
    return std::min(1024, ((thr + 63) / 64) * 64);
}

// Total cells of the ensemble, per-instance tiles and grid for a shape.
static int plan_shape(xf_ensemble* E, XfLaunch& L, bool commit) {
    // all instances share precision/ns/cpt/mode/block; bpt is per instance
    if (L.block < 64 || L.block > 1024 || L.block % 64)
        return xf_fail(XF_EINVAL, "block %d: a multiple of 64 in [64, 1024]", L.block);
    {
        int vg = 0, sp = 0;
        int kmax = xf_step_func_info(L, E->geom, &vg, &sp);
        if (kmax <= 0) return xf_fail(XF_EINVAL, "no stepper for this shape");
        if (L.block > kmax) return xf_fail(XF_EINVAL, "block %d exceeds the kernel's bound %d", L.block, kmax);
    }
    int grid = 0;
    std::vector<int> nblk(E->inst.size()), bpts(E->inst.size()), tiles(E->inst.size());
    int max_bpt_used = 1;
    L.ne = 1;                            // the per-energy LDS table holds the largest instance's bins
    for (auto& in : E->inst) L.ne = std::max(L.ne, in->t.ne);
    for (size_t i = 0; i < E->inst.size(); ++i) {
        XfInst* in = E->inst[i].get();
        int ne = in->t.ne;
        int bpt = std::min(L.bpt > 0 ? L.bpt : max_bpt(ne, L.cpt), max_bpt(ne, L.cpt));
        while (bpt > 1 && block_for(bpt, ne, L.cpt) > L.block) --bpt;
        if (block_for(bpt, ne, L.cpt) > L.block) return xf_fail(XF_EINVAL, "NE=%d does not fit block %d x cpt %d", ne, L.block, L.cpt);
        bpts[i] = bpt;
        tiles[i] = (in->t.nbeam + bpt - 1) / bpt;
        max_bpt_used = std::max(max_bpt_used, bpt);
    }
    XfLaunch Lq = L;
    Lq.bpt = max_bpt_used;              // LDS per workgroup depends on the largest tile
    Lq.wbeams = max_bpt_used;           // ... and in mode 4 on the records of one workgroup's tiles
    int maxb = 0, rc = 0;
    // in mode 4 the records grow as the workgroups get fewer and may lower
    // the occupancy: settle both (the counts only shrink)
    for (int pass = 0; pass < 4; ++pass) {
        rc = xf_step_max_blocks_geom(Lq, E->geom, &maxb);
        if (rc) return rc;
        grid = 0;
        int wbeams = max_bpt_used;
        if (L.mode != 2 && L.mode != 4) {
            for (size_t i = 0; i < E->inst.size(); ++i) nblk[i] = tiles[i];
        } else {
            // One workgroup per instance, then each further co-resident
            // workgroup to the instance whose workgroups carry the most tiles
            // (the step ends with the busiest workgroup; every tile costs one
            // pass over the block's cell slots).
            int budget = maxb;
            if (L.grid > 0) budget = std::min(budget, L.grid);
            std::priority_queue<std::pair<int, int>> busiest;     // (tiles per workgroup, instance)
            for (size_t i = 0; i < E->inst.size(); ++i) {
                nblk[i] = 1;
                if (tiles[i] > 1) busiest.push({tiles[i], (int)i});
            }
            for (int left = budget - (int)E->inst.size(); left > 0 && !busiest.empty(); --left) {
                const int i = busiest.top().second;
                busiest.pop();
                nblk[i] += 1;
                if (nblk[i] < tiles[i]) busiest.push({(tiles[i] + nblk[i] - 1) / nblk[i], i});
            }
            // as few workgroups as give the same busiest count, so that the
            // instance's workgroups carry equal counts where the tiles divide
            // (1000 tiles: 250 workgroups of 4 rather than 256 of 3 or 4),
            // and none waits in an exchange for a whole tile's work
            for (size_t i = 0; i < E->inst.size(); ++i) {
                const int per = (tiles[i] + nblk[i] - 1) / nblk[i];
                nblk[i] = (tiles[i] + per - 1) / per;
            }
        }
        for (size_t i = 0; i < E->inst.size(); ++i) {
            grid += nblk[i];
            wbeams = std::max(wbeams, (tiles[i] + nblk[i] - 1) / nblk[i] * bpts[i]);
        }
        if (L.mode == 4) {
            // the records of one workgroup's tiles, as many as LDS holds (the
            // kernel takes its tiles in chunks of wbeams / bpt)
            const int cap = xf_step_max_records(Lq, E->geom);
            if (cap < max_bpt_used)
                return xf_fail(XF_EINVAL, "mode 4: LDS holds %d beam records, a tile has %d", cap, max_bpt_used);
            wbeams = std::min(wbeams, cap);
        }
        if (L.mode != 4 || wbeams <= Lq.wbeams) break;
        Lq.wbeams = wbeams;
    }
    rc = xf_step_max_blocks_geom(Lq, E->geom, &maxb);
    if (rc) return rc;
    if (grid > maxb)
        return xf_fail(XF_EINVAL, "shape needs %d co-resident workgroups, device holds %d", grid, maxb);
    {
        // Every workgroup of the launch must be resident at once (the
        // exchanges wait for all of them), and the occupancy figure above
        // counts registers and LDS only.  A kernel that spills also needs
        // scratch memory for every wave in flight, which the runtime
        // provisions per dispatch against a per-device limit and may meet by
        // holding waves back: then some workgroups never start and the
        // others' exchange waits expire (XF_EBARRIER).  Shapes whose scratch
        // for the whole grid exceeds XF_COOP_SCRATCH_MAX are refused.
        int vg = 0, spill = 0;
        xf_step_func_info(Lq, E->geom, &vg, &spill);
        const double need = (double)spill * 64.0 * (double)grid * (double)(L.block / 64);
        if (spill > 0 && need > XF_COOP_SCRATCH_MAX)
            return xf_fail(XF_EINVAL, "shape spills %d B per lane: %.0f MiB of scratch for %d workgroups x %d threads "
                                      "exceeds the %.0f MiB a co-resident launch may use",
                           spill, need / 1048576.0, grid, L.block, XF_COOP_SCRATCH_MAX / 1048576.0);
    }
    L.wbeams = Lq.wbeams;
    if (!commit) {
        L.grid = grid;
        return XF_OK;
    }
    std::vector<XfInstDev> devs(E->inst.size());
    std::vector<int> blk_inst(grid);
    int b0 = 0;
    for (size_t i = 0; i < E->inst.size(); ++i) {
        XfInst* in = E->inst[i].get();
        rc = in->ensure_rows(nblk[i]);
        if (rc) return rc;
        in->dev.bpt = bpts[i];
        in->dev.ntiles = tiles[i];
        in->dev.tchunk = L.mode == 4 ? std::max(1, L.wbeams / bpts[i]) : 1;
        in->dev.blk0 = b0;
        in->dev.nblk = nblk[i];
        in->dev.launch_key = ((long long)nblk[i] << 44) | ((long long)L.block << 32) | ((long long)bpts[i] << 24) |
                             ((long long)(in->dev.tchunk & 0xFFFF) << 8) | ((long long)L.cpt << 4) | L.mode;
        for (int j = 0; j < nblk[i]; ++j) blk_inst[b0 + j] = (int)i;
        b0 += nblk[i];
        devs[i] = in->dev;
    }
    L.grid = grid;
    L.bpt = max_bpt_used;
    if (E->d_insts) (void)hipFree(E->d_insts);
    if (E->d_blk_inst) (void)hipFree(E->d_blk_inst);
    E->d_insts = nullptr;
    E->d_blk_inst = nullptr;
    if ((rc = dput(&E->d_insts, devs)) || (rc = dput(&E->d_blk_inst, blk_inst))) return rc;
    E->shadow = devs;
    for (auto& in : E->inst) in->rows_ok = false;       // a new workgroup map
    E->L = L;
    E->shaped = true;
    return XF_OK;
}

// Default shape: keep the state in registers when it fits, else stream.
static int choose_shape(xf_ensemble* E) {
    const xf_opts& o = E->o;
    XfLaunch base{};
    base.precision = o.precision;
    base.ns = E->inst[0]->ns;
    base.grid = o.grid;                 // streaming: at most this many workgroups (0: all co-resident)
    int ne_max = 1;
    for (auto& in : E->inst) ne_max = std::max(ne_max, in->t.ne);
    std::vector<int> modes, cpts;
    // streaming: mode 4 (mode 2's results, fewer set-ups), else mode 2 where
    // a tile's beam records do not fit LDS
    if (o.resident == 0) modes = {4, 2};
    else if (o.resident == 1) modes = {0, 1};
    else modes = {0, 1, 4, 2};
    if (o.cpt > 0) cpts = {o.cpt};
    else cpts = {1, 2, 4};
    for (int mode : modes) {
        for (int cpt : cpts) {
            XfLaunch L = base;
            L.mode = mode;
            L.cpt = cpt;
            if (!xf_step_has_shape(L, E->geom)) continue;
            int vg = 0, spill = 0;
            int kmax = xf_step_func_info(L, E->geom, &vg, &spill);
            if (kmax <= 0) continue;
            int bpt = max_bpt(ne_max, cpt);
            while (bpt > 1 && block_for(bpt, ne_max, cpt) > kmax) --bpt;
            L.block = o.block > 0 ? std::min(o.block, kmax) : block_for(bpt, ne_max, cpt);
            if (L.block > kmax) continue;
            // default shape without tuning: resident only when (almost)
            // spill-free (the kernel for this block)
            if (xf_step_func_info(L, E->geom, &vg, &spill) <= 0) continue;
            if (spill > 64 && mode != 2) continue;
            L.bpt = 0;
            if (plan_shape(E, L, false) == XF_OK) return plan_shape(E, L, true);
        }
    }
    return xf_fail(XF_EINVAL, "no launch shape fits the device: %s", xf_last_error());
}

// The descriptor of a one-instance launch on the device, copied only when it
// changed (stop predicates, the log, a trace).
static int ens_put_desc(xf_ensemble* E, const XfInstDev& d) {
    if (E->shadow.size() == 1 && std::memcmp(&d, &E->shadow[0], sizeof d) == 0) return XF_OK;
    HIPCHK(hipMemcpyAsync(E->d_insts, &d, sizeof d, hipMemcpyHostToDevice, E->stream));
    E->shadow.assign(1, d);
    return XF_OK;
}

// Clears the rows (and status) of every instance unless all of them can
// continue their generation count for a launch of max_steps steps.
static int ens_prepare_if_needed(xf_ensemble* E, int max_steps) {
    bool need = false;
    for (auto& in : E->inst)
        if (!in->rows_ok || (long long)in->gen0 + 3ll * max_steps + 2 > XF_MAX_EXCHANGES) need = true;
    if (!need) return XF_OK;
    int rc = xf_step_prepare(E->d_insts, (int)E->inst.size(), E->stream);
    if (rc) return rc;
    for (auto& in : E->inst) {
        in->rows_ok = true;
        in->gen0 = 0;
    }
    return XF_OK;
}

// psi0's moments for every instance unless the run scalars of all of them
// hold those of the current launch shape (d_insts / d_part: the launch's
// descriptors and a partial-sum buffer of XF_MOM_PARTS * 12 doubles per
// instance).
static int ens_moments_if_needed(xf_ensemble* E, const XfInstDev* d_insts, double* d_part) {
    bool need = false;
    for (auto& in : E->inst)
        if (in->m0_key != in->dev.launch_key) need = true;
    if (!need) return XF_OK;
    int nparts = 1;
    for (auto& in : E->inst) nparts = std::max(nparts, std::min(XF_MOM_PARTS, in->t.nbeam));
    int rc = xf_moments_launch(E->L, E->geom, d_insts, (int)E->inst.size(), nparts, d_part, E->stream);
    if (rc) return rc;
    for (auto& in : E->inst) in->m0_key = in->dev.launch_key;
    return XF_OK;
}
static int ens_part_buffer(xf_ensemble* E) {
    if (E->d_part) return XF_OK;
    return dalloc(&E->d_part, E->inst.size() * (size_t)XF_MOM_PARTS * 12);
}

// After a launch drained: every instance's outcome from its pinned copy.
static void ens_collect(xf_ensemble* E) {
    for (auto& in : E->inst) {
        in->last = *in->h_rs;
        in->gen0 = in->last.gen0;
        in->m0_key = in->last.m0_ok ? in->last.m0_key : 0;
        if (in->last.status) in->rows_ok = false;
    }
}

static int ens_launch(xf_ensemble* E, int max_steps, xf_step_rec* log, double* ms) {
    if (!E->shaped) {
        int rc = choose_shape(E);
        if (rc) return rc;
    }
    if (log && E->inst.size() == 1) {
        if (max_steps > E->log_cap) {
            if (E->d_log) (void)hipFree(E->d_log);
            E->d_log = nullptr;
            int rc = dalloc(&E->d_log, (size_t)max_steps);
            if (rc) return rc;
            E->log_cap = max_steps;
        }
    }
    if (E->inst.size() == 1) {
        XfInstDev d = E->inst[0]->dev;
        d.log = log ? E->d_log : nullptr;
        int rc = ens_put_desc(E, d);
        if (rc) return rc;
    }
    if (3ll * max_steps + 1 > XF_MAX_EXCHANGES)
        return xf_fail(XF_EINVAL, "%d steps in one launch exceed the exchange tags", max_steps);
    int rc0 = ens_prepare_if_needed(E, max_steps);
    if (rc0) return rc0;
    if ((rc0 = ens_part_buffer(E)) || (rc0 = ens_moments_if_needed(E, E->d_insts, E->d_part))) return rc0;
    HIPCHK(hipEventRecord(E->ev0, E->stream));
    int rc = xf_step_launch_geom(E->L, E->geom, E->d_insts, E->d_blk_inst, max_steps, E->stream);
    if (rc) return rc;
    HIPCHK(hipEventRecord(E->ev1, E->stream));
    HIPCHK(hipStreamSynchronize(E->stream));
    ens_collect(E);
    float f = 0;
    HIPCHK(hipEventElapsedTime(&f, E->ev0, E->ev1));
    if (ms) *ms = f;
    if (log && E->inst.size() == 1) {
        const long long n = E->inst[0]->last.steps_call;
        if (n > 0) HIPCHK(hipMemcpy(log, E->d_log, (size_t)n * sizeof(xf_step_rec), hipMemcpyDeviceToHost));
    }
    return XF_OK;
}

// `launches` launches of up to max_steps steps from one instantiated graph:
// the rows and status are cleared once before it, and the launches continue
// one generation count (a failed launch leaves its status, which ends the
// rest).  Per-instance run scalars before the graph go to *before.
static int ens_run_graph(xf_ensemble* E, int launches, int max_steps, std::vector<XfRunScalars>* before, double* ms) {
    if (launches < 1 || max_steps < 1) return xf_fail(XF_EINVAL, "graph of %d launches x %d steps", launches, max_steps);
    if ((long long)launches * (3ll * max_steps + 1) + 1 > XF_MAX_EXCHANGES)
        return xf_fail(XF_EINVAL, "%d launches of %d steps exceed the exchange tags", launches, max_steps);
    if (!E->shaped) {
        int rc = choose_shape(E);
        if (rc) return rc;
    }
    if (E->inst.size() == 1) {
        XfInstDev d = E->inst[0]->dev;
        d.log = nullptr;
        int rc = ens_put_desc(E, d);
        if (rc) return rc;
    }
    before->resize(E->inst.size());
    for (size_t i = 0; i < E->inst.size(); ++i) {
        int rc = E->inst[i]->scalars(&(*before)[i]);
        if (rc) return rc;
    }
    if (!E->gexec || E->g_launches != launches || E->g_steps != max_steps ||
        std::memcmp(&E->g_L, &E->L, sizeof(XfLaunch)) != 0 || E->g_insts != E->d_insts || E->g_blk != E->d_blk_inst) {
        if (E->gexec) (void)hipGraphExecDestroy(E->gexec);
        E->gexec = nullptr;
        int rc = xf_step_build_graph(E->L, E->geom, E->d_insts, E->d_blk_inst, (int)E->inst.size(), max_steps, launches,
                                     &E->gexec);
        if (rc) return rc;
        E->g_launches = launches;
        E->g_steps = max_steps;
        E->g_L = E->L;
        E->g_insts = E->d_insts;
        E->g_blk = E->d_blk_inst;
    }
    int rc = xf_step_prepare(E->d_insts, (int)E->inst.size(), E->stream);
    if (rc) return rc;
    for (auto& in : E->inst) {
        in->rows_ok = true;
        in->gen0 = 0;
    }
    if ((rc = ens_part_buffer(E)) || (rc = ens_moments_if_needed(E, E->d_insts, E->d_part))) return rc;
    HIPCHK(hipEventRecord(E->ev0, E->stream));
    HIPCHK(hipGraphLaunch(E->gexec, E->stream));
    HIPCHK(hipEventRecord(E->ev1, E->stream));
    HIPCHK(hipStreamSynchronize(E->stream));
    // the last launch's copy; the graph's own clears restarted the count
    // before each launch, so the rows agree with its gen0
    ens_collect(E);
    float f = 0;
    HIPCHK(hipEventElapsedTime(&f, E->ev0, E->ev1));
    if (ms) *ms = f;
    return XF_OK;
}

static int ens_create(const xf_params* p, int n, const xf_opts* o, xf_ensemble** out) {
    if (!p || n <= 0 || !out) return xf_fail(XF_EINVAL, "xf_ensemble_create: bad arguments");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return xf_fail(XF_ENODEV, "no HIP device");
    std::unique_ptr<xf_ensemble> E(new xf_ensemble());
    if (o) E->o = *o;
    if (E->o.precision < XF_FP64 || E->o.precision > XF_FP32) return xf_fail(XF_EINVAL, "precision %d", E->o.precision);
    int ns = E->o.species == 4 ? 4 : 2;
    if (ns == 4 && E->o.precision == XF_FP32)
        return xf_fail(XF_EINVAL, "explicit 4-species states are stepped in FP64 or mixed precision: the "
                                  "FP32-state stepper takes SU(2)-partner states (species 2)");
    int ren = E->o.renorm;
    if (ren == -1) ren = (E->o.precision != XF_FP64) ? 1 : 0;     // automatic
    else if (ren < 0) ren = 0;
    E->geom = p[0].module;
    for (int i = 0; i < n; ++i) {
        if (p[i].module != E->geom) return xf_fail(XF_EINVAL, "all instances of an ensemble share the geometry module");
        std::unique_ptr<XfInst> in(new XfInst());
        int rc = in->init(&p[i], E->o.precision, ns, ren);
        if (rc) return rc;
        E->inst.push_back(std::move(in));
    }
    HIPCHK(hipStreamCreateWithFlags(&E->stream, hipStreamNonBlocking));
    HIPCHK(hipEventCreate(&E->ev0));
    HIPCHK(hipEventCreate(&E->ev1));
    *out = E.release();
    return XF_OK;
}

// xf_run_info from run scalars (of the device, or the copy a launch left).
static int fill_info_from(const XfRunScalars& rs, xf_run_info* info, double ms) {
    if (info) {
        info->steps = rs.steps_call;
        info->accepted = rs.acc_call;
        info->total_steps = rs.total;
        info->total_accepted = rs.accepted;
        info->r = rs.r;
        info->dr = rs.dr;
        info->last_err = rs.last_err;
        info->last_dr = rs.last_dr;
        info->kernel_ms = ms;
        info->stop_reason = rs.status ? 3 : rs.stop_reason;
    }
    if (rs.status == XF_EBARRIER && rs.diag_where)
        return xf_fail(rs.status, "stepper status %d (an exchange wait expired: workgroup %d, %s, waited for generation "
                                  "tag %u; %d rows missing, the first, row %d, holds tag bits %u)", rs.status, rs.diag_wg,
                       rs.diag_where == 1 ? "first check" : "fetch", rs.diag_tag, rs.diag_missing, rs.diag_row, rs.diag_seen);
    if (rs.status)
        return xf_fail(rs.status, "stepper status %d (%s)", rs.status,
                       rs.status == XF_EBARRIER ? "an exchange wait expired"
                       : rs.status == XF_ESTATE ? "no moments of psi0 for this launch shape" : "launch error");
    return XF_OK;
}
static int fill_info(XfInst* in, xf_run_info* info, double ms) {
    XfRunScalars rs;
    int rc = in->scalars(&rs);
    if (rc) return rc;
    return fill_info_from(rs, info, ms);
}

// Times every instantiated (mode, cells per lane, beams per tile) on the
// ensemble's current states for `steps` attempted steps and keeps the
// fastest; every instance's state and run scalars are restored afterwards.
static int ens_autotune(xf_ensemble* E, int steps, xf_opts* best, double* best_ms_per_step) {
    struct Saved { void* psi = nullptr; size_t bytes = 0; XfRunScalars rs{}; };
    std::vector<Saved> sv(E->inst.size());
    int rc = XF_OK;
    int ne_max = 1;
    for (size_t i = 0; i < E->inst.size(); ++i) {
        XfInst* in = E->inst[i].get();
        ne_max = std::max(ne_max, in->t.ne);
        sv[i].bytes = (size_t)in->ns * 4 * in->ncells() * in->elem;
        if ((rc = in->scalars(&sv[i].rs))) break;
        if (hipMalloc(&sv[i].psi, sv[i].bytes) != hipSuccess) { rc = xf_fail(XF_ENOMEM, "autotune: state copy"); break; }
        if (hipMemcpy(sv[i].psi, in->d_psi[sv[i].rs.cur], sv[i].bytes, hipMemcpyDeviceToDevice) != hipSuccess) {
            rc = xf_fail(XF_EHIP, "autotune: state copy");
            break;
        }
    }
    auto restore = [&]() -> int {
        for (size_t i = 0; i < E->inst.size(); ++i) {
            XfInst* in = E->inst[i].get();
            if (!sv[i].psi) continue;
            HIPCHK(hipMemcpy(in->d_psi[0], sv[i].psi, sv[i].bytes, hipMemcpyDeviceToDevice));
            XfRunScalars rs = sv[i].rs;
            rs.cur = 0;
            HIPCHK(hipMemcpy(in->d_rs, &rs, sizeof rs, hipMemcpyHostToDevice));
            in->rows_ok = false;     // the trial launches moved the rows past rs.gen0
            in->m0_key = rs.m0_ok ? rs.m0_key : 0;
            in->last = rs;
        }
        return XF_OK;
    };
    double best_t = 1e300;
    XfLaunch bestL{};
    bool found = false;
    const XfLaunch prev = E->L;
    const bool had_shape = E->shaped;
    for (int mode : {0, 1, 2, 4})
        for (int cpt : {1, 2, 4}) {
            if (rc != XF_OK) break;
            XfLaunch L{};
            L.precision = E->o.precision;
            L.ns = E->inst[0]->ns;
            L.mode = mode;
            L.cpt = cpt;
            if (!xf_step_has_shape(L, E->geom)) continue;
            // beams per tile: the eight largest that fit the kernel's thread
            // bound (they decide how evenly the beams spread over the
            // co-resident workgroups), then halvings
            int bmax = max_bpt(ne_max, cpt);
            {
                int vg = 0, spill = 0;
                const int kmax = xf_step_func_info(L, E->geom, &vg, &spill);
                while (bmax > 1 && block_for(bmax, ne_max, cpt) > kmax) --bmax;
            }
            // (bpt cap, block).  An ensemble also tries cap 0: the block that
            // fits the largest instance's tile, and in it every instance's
            // largest tile.  A cap sized for the largest NE would leave
            // instances of small NE with tiles that fill a few lanes.
            std::vector<std::pair<int, int>> cands;
            if (E->inst.size() > 1) cands.push_back({0, block_for(bmax, ne_max, cpt)});
            for (int b = bmax; b >= 1 && b > bmax - 8; --b) cands.push_back({b, block_for(b, ne_max, cpt)});
            for (int b = (bmax - 8) / 2; b >= 1; b /= 2) cands.push_back({b, block_for(b, ne_max, cpt)});
            for (const auto& cand : cands) {
                L.bpt = cand.first;
                L.block = cand.second;
                L.grid = 0;
                int vg = 0, spill = 0;
                int kmax = xf_step_func_info(L, E->geom, &vg, &spill);
                if (kmax <= 0 || L.block > kmax) continue;
                // a shape the planner refuses (does not fit the device, or
                // would need too much scratch) is not a candidate; a shape
                // that was launched and failed ends the tuning with its error
                XfLaunch Lp = L;
                if (plan_shape(E, Lp, true) != XF_OK) continue;
                if ((rc = restore())) break;
                double ms = 0;
                if ((rc = ens_launch(E, 2, nullptr, &ms)) || (rc = ens_launch(E, steps, nullptr, &ms))) {
                    rc = xf_fail(rc, "autotune: shape mode %d, cpt %d, block %d, bpt %d failed: %s", E->L.mode,
                                 E->L.cpt, E->L.block, E->L.bpt, xf_last_error());
                    break;
                }
                long long done = 0;
                for (auto& in : E->inst) {
                    XfRunScalars r2;
                    if ((rc = in->scalars(&r2))) break;
                    if (r2.status) {
                        fill_info_from(r2, nullptr, 0.);
                        const std::string why = xf_last_error();
                        rc = xf_fail(r2.status, "autotune: shape mode %d, cpt %d, block %d, bpt %d, grid %d ended with "
                                                "status %d: %s", E->L.mode, E->L.cpt, E->L.block, E->L.bpt, E->L.grid,
                                     r2.status, why.c_str());
                        break;
                    }
                    done = std::max<long long>(done, r2.steps_call);
                }
                if (rc) break;
                if (done <= 0) continue;
                const double per = ms / (double)done;
                if (per < best_t) {
                    best_t = per;
                    bestL = E->L;
                    found = true;
                }
            }
        }
    int rc2 = restore();
    for (auto& x : sv)
        if (x.psi) (void)hipFree(x.psi);
    if (rc) return rc;
    if (rc2) return rc2;
    if (!found) {
        if (had_shape) { XfLaunch L = prev; plan_shape(E, L, true); }
        return xf_fail(XF_EINVAL, "autotune: no shape ran");
    }
    XfLaunch L = bestL;
    rc = plan_shape(E, L, true);
    if (rc) return rc;
    if (best) {
        *best = E->o;
        best->species = E->inst[0]->ns;
        best->block = E->L.block;
        best->cpt = E->L.cpt;
        best->resident = (E->L.mode == 0 || E->L.mode == 1) ? 1 : 0;
        best->mode = E->L.mode;
        best->grid = E->L.grid;
    }
    if (best_ms_per_step) *best_ms_per_step = best_t;
    return XF_OK;
}

// ================================================================ C-ABI
extern "C" {

int xf_device_count(int* n) {
    int d = 0;
    if (hipGetDeviceCount(&d) != hipSuccess) d = 0;
    if (n) *n = d;
    return XF_OK;
}

int xf_create(const xf_params* p, const xf_opts* o, xf_ctx** out) {
    if (!p || !out) return xf_fail(XF_EINVAL, "xf_create: null argument");
    std::unique_ptr<xf_ctx> c(new xf_ctx());
    c->p = *p;
    if (o) c->o = *o;
    else { c->o.renorm = -1; c->o.resident = -1; }
    int rc = xf_build_tables(p, &c->t);
    if (rc) return rc;
    rc = ens_create(p, 1, &c->o, &c->ens);
    if (rc) return rc;
    rc = xf_state_init_pure(c.get());
    if (rc) return rc;
    *out = c.release();
    return XF_OK;
}

int xf_destroy(xf_ctx* ctx) {
    if (!ctx) return XF_OK;
    delete ctx->ens;
    xf_phy_destroy(ctx->phy);
    delete ctx;
    return XF_OK;
}

int xf_beam_len(const xf_ctx* ctx) { return ctx ? ctx->t.nbeam * 4 : 0; }
int xf_get_dim(const xf_ctx* ctx) {
    if (!ctx) return 0;
    return ctx->t.module == XF_MAA ? 6 : ctx->t.module == XF_SA ? 4 : 5;
}
int xf_count_dim(const xf_ctx* ctx, size_t* count) {
    if (!ctx || !count) return xf_fail(XF_EINVAL, "xf_count_dim: null argument");
    // [r, theta, (phi,) prtcl, comp, ebin] (Phy_MA.cpp:253-258, Phy_MAA.cpp:269-275);
    // SA: [r, prtcl, comp, ebin] (Phy_SA.cpp:74-80)
    int k = 0;
    count[k++] = 1;
    if (ctx->t.module != XF_SA) count[k++] = ctx->t.nth;
    if (ctx->t.module == XF_MAA) count[k++] = ctx->t.nph;
    count[k++] = 4;
    count[k++] = 4;
    count[k++] = ctx->t.ne;
    return XF_OK;
}

int xf_get_spectra(const xf_ctx* ctx, double* fv, double* hvac, double* L_E, double* Ev) {
    if (!ctx) return xf_fail(XF_EINVAL, "null ctx");
    if (fv) std::memcpy(fv, ctx->t.fv.data(), ctx->t.fv.size() * 8);
    if (hvac) std::memcpy(hvac, ctx->t.hvac.data(), ctx->t.hvac.size() * 8);
    for (int s = 0; s < 4; ++s) {
        if (L_E) L_E[s] = ctx->t.L_E[s];
        if (Ev) Ev[s] = ctx->t.Ev[s];
    }
    return XF_OK;
}

int xf_get_geometry(const xf_ctx* ctx, double* cos2R, double* coef) {
    if (!ctx) return xf_fail(XF_EINVAL, "null ctx");
    if (cos2R) std::memcpy(cos2R, ctx->t.cos2R.data(), ctx->t.cos2R.size() * 8);
    if (coef) *coef = ctx->t.integ_coef;
    return XF_OK;
}

int xf_get_hm(const xf_ctx* ctx, double r, double* hm) {
    if (!ctx || !hm) return xf_fail(XF_EINVAL, "null argument");
    *hm = xf_host_hm(&ctx->p, &ctx->t, r);
    return XF_OK;
}

// ---- fused stepper
int xf_state_upload(xf_ctx* ctx, const double* psi, double r, double dr) {
    if (!ctx || !psi) return xf_fail(XF_EINVAL, "xf_state_upload: null argument");
    XfInst* in = ctx->ens->inst[0].get();
    int rc = in->upload(psi, r, dr);
    if (rc == XF_ESTATE && ctx->o.species == 0 && ctx->o.precision != XF_FP32) {
        // not partners: fall back to the explicit 4-species state (FP64, mixed)
        xf_opts o4 = ctx->o;
        o4.species = 4;
        xf_ensemble* E = nullptr;
        rc = ens_create(&ctx->p, 1, &o4, &E);
        if (rc) return rc;
        delete ctx->ens;
        ctx->ens = E;
        return E->inst[0]->upload(psi, r, dr);
    }
    return rc;
}

int xf_state_init_pure(xf_ctx* ctx) {
    if (!ctx) return xf_fail(XF_EINVAL, "null ctx");
    const int ne = ctx->t.ne;
    std::vector<double> psi((size_t)ctx->t.nbeam * 16 * ne, 0.);
    for (int b = 0; b < ctx->t.nbeam; ++b)
        for (int s = 0; s < 4; ++s)
            for (int e = 0; e < ne; ++e) {
                // NBeam::NBeam(int): a = 1 for nu_e / anti-nu_e, b = 1 for nu_x / anti-nu_x
                psi[(((size_t)b * 4 + s) * 4 + (s < 2 ? 0 : 2)) * ne + e] = 1.;
            }
    return ctx->ens->inst[0]->upload(psi.data(), ctx->p.R0, ctx->p.dr);
}

int xf_state_download(xf_ctx* ctx, double* psi, double* r, double* dr) {
    if (!ctx) return xf_fail(XF_EINVAL, "null ctx");
    return ctx->ens->inst[0]->download(psi, r, dr);
}

int xf_state_counters(const xf_ctx* ctx, int64_t* total, int64_t* accepted) {
    if (!ctx) return xf_fail(XF_EINVAL, "null ctx");
    XfRunScalars rs;
    int rc = ctx->ens->inst[0]->scalars(&rs);
    if (rc) return rc;
    if (total) *total = rs.total;
    if (accepted) *accepted = rs.accepted;
    return XF_OK;
}

int xf_state_species(const xf_ctx* ctx) { return ctx ? ctx->ens->inst[0]->ns : 0; }

// Rn of one instance; a run that ended on the old Rn (dr = 0) restarts with
// the length of its last accepted step.  The descriptor copy on the device is
// refreshed: xf_run_steps copies it before every launch, an ensemble here.
static int inst_set_rn(xf_ensemble* E, size_t i, double rn) {
    if (!(rn > 0.) || !std::isfinite(rn)) return xf_fail(XF_EINVAL, "Rn = %g", rn);
    XfInst* in = E->inst[i].get();
    XfRunScalars rs;
    int rc = in->scalars(&rs);
    if (rc) return rc;
    if (rs.dr == 0. && rn > rs.r && rs.last_dr > 0.) {
        rs.dr = std::min(rs.last_dr, rn - rs.r);
        HIPCHK(hipMemcpy(in->d_rs, &rs, sizeof rs, hipMemcpyHostToDevice));
    }
    in->p.Rn = rn;
    in->dev.Rn = rn;
    if (E->shaped && E->d_insts) {
        HIPCHK(hipMemcpy(E->d_insts + i, &in->dev, sizeof(XfInstDev), hipMemcpyHostToDevice));
        if (i < E->shadow.size()) E->shadow[i] = in->dev;
    }
    return XF_OK;
}

int xf_set_rn(xf_ctx* ctx, double rn) {
    if (!ctx) return xf_fail(XF_EINVAL, "xf_set_rn: null ctx");
    int rc = inst_set_rn(ctx->ens, 0, rn);
    if (rc == XF_OK) ctx->p.Rn = rn;
    return rc;
}

int xf_run_steps(xf_ctx* ctx, int max_steps, const xf_stop* stop, xf_step_rec* log, xf_run_info* info) {
    if (!ctx || max_steps < 0) return xf_fail(XF_EINVAL, "xf_run_steps: bad arguments");
    xf_ensemble* E = ctx->ens;
    XfInst* in = E->inst[0].get();
    in->dev.dump_rstep = stop ? stop->dump_rstep : 0.;
    in->dev.dump_last_r = stop ? stop->dump_last_r : 0.;
    in->dev.dump_itrstep = stop ? stop->dump_itrstep : 0;
    double ms = 0;
    int rc = ens_launch(E, max_steps, log, &ms);     // refreshes the descriptor copy where it changed
    if (rc) return rc;
    return fill_info_from(in->last, info, ms);
}

// xf_run_steps with the descriptor, the rows and the log in the static
// scratch array of xf_noalloc.hip: no allocation on this path (the launch
// shape, the stream and the events were made at set-up).
int xf_run_steps_noalloc(xf_ctx* ctx, int max_steps, const xf_stop* stop, xf_step_rec* log, xf_run_info* info) {
    if (!ctx || max_steps < 0) return xf_fail(XF_EINVAL, "xf_run_steps_noalloc: bad arguments");
    xf_ensemble* E = ctx->ens;
    if (E->inst.size() != 1) return xf_fail(XF_EINVAL, "xf_run_steps_noalloc: one instance per context");
    XfInst* in = E->inst[0].get();
    if (!E->shaped)
        return xf_fail(XF_EINVAL, "xf_run_steps_noalloc: no launch shape; choose one at set-up "
                                  "(xf_get_launch, xf_set_launch or xf_autotune)");
    if (in->dev.nblk > XF_NOALLOC_MAX_WORKGROUPS || E->L.grid > XF_NOALLOC_MAX_WORKGROUPS)
        return xf_fail(XF_EINVAL, "xf_run_steps_noalloc: the launch has %d workgroups, the scratch array holds "
                                  "exchange rows for %d", E->L.grid, XF_NOALLOC_MAX_WORKGROUPS);
    if (log && max_steps > XF_NOALLOC_MAX_LOG)
        return xf_fail(XF_EINVAL, "xf_run_steps_noalloc: a log of %d attempts exceeds the scratch array's %d",
                       max_steps, XF_NOALLOC_MAX_LOG);
    if (3ll * max_steps + 1 > XF_MAX_EXCHANGES)
        return xf_fail(XF_EINVAL, "%d steps in one launch exceed the exchange tags", max_steps);
    XfNoAllocArea A{};
    int rc = xf_noalloc_area(&A);
    if (rc) return rc;
    XfInstDev d = in->dev;
    d.rows = A.rows;
    d.log = log ? A.log : nullptr;
    d.trace = nullptr;
    d.trace_steps = 0;
    d.dump_rstep = stop ? stop->dump_rstep : 0.;
    d.dump_last_r = stop ? stop->dump_last_r : 0.;
    d.dump_itrstep = stop ? stop->dump_itrstep : 0;
    if ((rc = xf_noalloc_setup(d, A, E->L.grid, E->stream))) return rc;
    // the shared rows start from zero (and the count from 0) on every call;
    // the instance's own rows no longer follow its count
    in->rows_ok = false;
    if ((rc = xf_step_prepare(A.desc, 1, E->stream))) return rc;
    if ((rc = ens_moments_if_needed(E, A.desc, A.part))) return rc;
    HIPCHK(hipEventRecord(E->ev0, E->stream));
    if ((rc = xf_step_launch_geom(E->L, E->geom, A.desc, A.blk, max_steps, E->stream))) return rc;
    HIPCHK(hipEventRecord(E->ev1, E->stream));
    HIPCHK(hipStreamSynchronize(E->stream));
    in->last = *in->h_rs;
    in->gen0 = in->last.gen0;
    in->m0_key = in->last.m0_ok ? in->last.m0_key : 0;
    float f = 0;
    HIPCHK(hipEventElapsedTime(&f, E->ev0, E->ev1));
    if (log && in->last.steps_call > 0)
        HIPCHK(hipMemcpy(log, A.log, (size_t)in->last.steps_call * sizeof(xf_step_rec), hipMemcpyDeviceToHost));
    return fill_info_from(in->last, info, f);
}

int xf_run_graph(xf_ctx* ctx, int launches, int max_steps, xf_run_info* info) {
    if (!ctx) return xf_fail(XF_EINVAL, "xf_run_graph: null ctx");
    xf_ensemble* E = ctx->ens;
    XfInst* in = E->inst[0].get();
    in->dev.dump_rstep = 0.;
    in->dev.dump_last_r = 0.;
    in->dev.dump_itrstep = 0;
    std::vector<XfRunScalars> before;
    double ms = 0;
    int rc = ens_run_graph(E, launches, max_steps, &before, &ms);
    if (rc) return rc;
    rc = fill_info(in, info, ms);
    if (info) {
        info->steps = info->total_steps - before[0].total;
        info->accepted = info->total_accepted - before[0].accepted;
    }
    return rc;
}

int xf_get_launch(const xf_ctx* ctx, xf_opts* out, int* tiles, int* beams_per_tile) {
    if (!ctx) return xf_fail(XF_EINVAL, "null ctx");
    xf_ensemble* E = ctx->ens;
    if (!E->shaped) {
        int rc = choose_shape(E);
        if (rc) return rc;
    }
    if (out) {
        *out = E->o;
        out->species = E->inst[0]->ns;
        out->block = E->L.block;
        out->cpt = E->L.cpt;
        out->resident = (E->L.mode == 0 || E->L.mode == 1) ? 1 : 0;
        out->mode = E->L.mode;      // 0 keep, 1 recompute, 2 / 4 stream
        out->grid = E->L.grid;
        out->renorm = E->inst[0]->renorm;
    }
    if (tiles) *tiles = E->inst[0]->dev.ntiles;
    if (beams_per_tile) *beams_per_tile = E->inst[0]->dev.bpt;
    return XF_OK;
}

// Set an explicit launch shape (mode: 0 keep, 1 recompute, 2 stream).
int xf_set_launch(xf_ctx* ctx, int mode, int cpt, int block, int bpt) {
    if (!ctx) return xf_fail(XF_EINVAL, "null ctx");
    if (mode < 0 || mode > 4 || mode == 3)
        return xf_fail(XF_EINVAL, "mode %d: 0 keep, 1 recompute, 2 stream, 4 stream with the records in LDS", mode);
    if (cpt != 1 && cpt != 2 && cpt != 4) return xf_fail(XF_EINVAL, "cpt %d: 1, 2 or 4", cpt);
    if (block < 64 || block > 1024 || block % 64) return xf_fail(XF_EINVAL, "block %d: a multiple of 64 in [64, 1024]", block);
    if (bpt < 1 || bpt > 128) return xf_fail(XF_EINVAL, "bpt %d: 1..128 beams per tile", bpt);
    xf_ensemble* E = ctx->ens;
    XfLaunch L{};
    L.precision = E->o.precision;
    L.ns = E->inst[0]->ns;
    L.mode = mode;
    L.cpt = cpt;
    L.block = block;
    L.bpt = bpt;
    L.grid = E->o.grid;                 // streaming: at most this many workgroups (0: all co-resident)
    if (!xf_step_has_shape(L, E->geom)) return xf_fail(XF_EINVAL, "shape not instantiated");
    return plan_shape(E, L, true);
}

int xf_launch_info(const xf_ctx* ctx, int* vgprs, int* spill, int* max_blocks, int* lds_bytes) {
    if (!ctx || !ctx->ens->shaped) return xf_fail(XF_EINVAL, "no launch shape yet");
    const xf_ensemble* E = ctx->ens;
    int rc = xf_step_func_info(E->L, E->geom, vgprs, spill);
    if (rc < 0) return rc;
    if (max_blocks) {
        rc = xf_step_max_blocks_geom(E->L, E->geom, max_blocks);
        if (rc) return rc;
    }
    if (lds_bytes) *lds_bytes = (int)xf_step_smem_bytes(E->L, E->geom);
    return XF_OK;
}

// Per-phase real-time stamps (100 MHz counter) of workgroup 0 for the first
// `steps` steps of the next xf_run_steps; 16 slots per step: [0] step start,
// [1..4] pass A cells done / partial published / arrived / totals known,
// [5..8] the same for pass B, [9..12] pass C, [13] decision.
int xf_trace_next(xf_ctx* ctx, long long* out, int steps) {
    if (!ctx || steps < 0) return xf_fail(XF_EINVAL, "bad arguments");
    XfInst* in = ctx->ens->inst[0].get();
    static thread_local long long* d_tr = nullptr;
    static thread_local int cap = 0;
    if (!out || steps == 0) {
        in->dev.trace = nullptr;
        in->dev.trace_steps = 0;
        return XF_OK;
    }
    if (steps > cap) {
        if (d_tr) (void)hipFree(d_tr);
        d_tr = nullptr;
        int rc = dalloc(&d_tr, (size_t)steps * 16);
        if (rc) return rc;
        cap = steps;
    }
    HIPCHK(hipMemset(d_tr, 0, (size_t)steps * 16 * sizeof(long long)));
    in->dev.trace = d_tr;
    in->dev.trace_steps = steps;
    return XF_OK;
}
int xf_trace_fetch(xf_ctx* ctx, long long* out, int steps) {
    if (!ctx || !out) return xf_fail(XF_EINVAL, "bad arguments");
    XfInst* in = ctx->ens->inst[0].get();
    if (!in->dev.trace) return xf_fail(XF_EINVAL, "no trace armed");
    HIPCHK(hipMemcpy(out, in->dev.trace, (size_t)std::min(steps, in->dev.trace_steps) * 16 * sizeof(long long), hipMemcpyDeviceToHost));
    in->dev.trace = nullptr;
    in->dev.trace_steps = 0;
    return XF_OK;
}

int xf_autotune(xf_ctx* ctx, int steps, xf_opts* best, double* best_ms_per_step) {
    if (!ctx || steps <= 0) return xf_fail(XF_EINVAL, "xf_autotune: bad arguments");
    return ens_autotune(ctx->ens, steps, best, best_ms_per_step);
}

// ---- ensemble
int xf_ensemble_create(const xf_params* p, int n, const xf_opts* o, xf_ensemble** out) {
    xf_opts oo{};
    if (o) oo = *o;
    else { oo.renorm = -1; oo.resident = -1; }
    return ens_create(p, n, &oo, out);
}
int xf_ensemble_destroy(xf_ensemble* ens) {
    delete ens;
    return XF_OK;
}
int xf_ensemble_upload(xf_ensemble* ens, int i, const double* psi, double r, double dr) {
    if (!ens || i < 0 || i >= (int)ens->inst.size() || !psi) return xf_fail(XF_EINVAL, "xf_ensemble_upload: bad arguments");
    return ens->inst[i]->upload(psi, r, dr);
}
int xf_ensemble_init_pure(xf_ensemble* ens, int i) {
    if (!ens || i < 0 || i >= (int)ens->inst.size()) return xf_fail(XF_EINVAL, "bad instance");
    XfInst* in = ens->inst[i].get();
    const int ne = in->t.ne;
    std::vector<double> psi((size_t)in->t.nbeam * 16 * ne, 0.);
    for (int b = 0; b < in->t.nbeam; ++b)
        for (int s = 0; s < 4; ++s)
            for (int e = 0; e < ne; ++e) psi[(((size_t)b * 4 + s) * 4 + (s < 2 ? 0 : 2)) * ne + e] = 1.;
    return in->upload(psi.data(), in->p.R0, in->p.dr);
}
int xf_ensemble_download(xf_ensemble* ens, int i, double* psi, double* r, double* dr) {
    if (!ens || i < 0 || i >= (int)ens->inst.size()) return xf_fail(XF_EINVAL, "bad instance");
    return ens->inst[i]->download(psi, r, dr);
}
int xf_ensemble_autotune(xf_ensemble* ens, int steps, xf_opts* best, double* best_ms_per_step) {
    if (!ens || steps <= 0) return xf_fail(XF_EINVAL, "xf_ensemble_autotune: bad arguments");
    return ens_autotune(ens, steps, best, best_ms_per_step);
}

int xf_ensemble_run_steps(xf_ensemble* ens, int max_steps, xf_run_info* info) {
    if (!ens || max_steps < 0) return xf_fail(XF_EINVAL, "bad arguments");
    double ms = 0;
    int rc = ens_launch(ens, max_steps, nullptr, &ms);
    if (rc) return rc;
    int first_err = XF_OK;
    for (size_t i = 0; i < ens->inst.size(); ++i) {
        rc = fill_info_from(ens->inst[i]->last, info ? info + i : nullptr, ms);
        if (rc && !first_err) first_err = rc;
    }
    return first_err;
}
int xf_ensemble_run_graph(xf_ensemble* ens, int launches, int max_steps, xf_run_info* info) {
    if (!ens) return xf_fail(XF_EINVAL, "bad arguments");
    std::vector<XfRunScalars> before;
    double ms = 0;
    int rc = ens_run_graph(ens, launches, max_steps, &before, &ms);
    if (rc) return rc;
    int first_err = XF_OK;
    for (size_t i = 0; i < ens->inst.size(); ++i) {
        xf_run_info* x = info ? info + i : nullptr;
        rc = fill_info(ens->inst[i].get(), x, ms);
        if (x) {
            x->steps = x->total_steps - before[i].total;
            x->accepted = x->total_accepted - before[i].accepted;
        }
        if (rc && !first_err) first_err = rc;
    }
    return first_err;
}
int xf_ensemble_cells(const xf_ensemble* ens, int64_t* cells) {
    if (!ens || !cells) return xf_fail(XF_EINVAL, "bad arguments");
    int64_t c = 0;
    for (auto& in : ens->inst) c += in->ncells();
    *cells = c;
    return XF_OK;
}
int xf_ensemble_set_rn(xf_ensemble* ens, int i, double rn) {
    if (!ens || i < 0 || i >= (int)ens->inst.size()) return xf_fail(XF_EINVAL, "xf_ensemble_set_rn: bad instance");
    return inst_set_rn(ens, (size_t)i, rn);
}
int xf_ensemble_launch(const xf_ensemble* ens, int* mode, int* cpt, int* block, int* grid) {
    if (!ens || !ens->shaped) return xf_fail(XF_EINVAL, "no launch shape yet");
    if (mode) *mode = ens->L.mode;
    if (cpt) *cpt = ens->L.cpt;
    if (block) *block = ens->L.block;
    if (grid) *grid = ens->L.grid;
    return XF_OK;
}

}  // extern "C"

// access for the Phy mirror
XfPhy*& xf_ctx_phy(xf_ctx* ctx) { return ctx->phy; }
const xf_params& xf_ctx_params(const xf_ctx* ctx) { return ctx->p; }
const XfTables& xf_ctx_tables(const xf_ctx* ctx) { return ctx->t; }
