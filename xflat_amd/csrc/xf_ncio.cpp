// xf_ncio.cpp -- NetCDF classic-format codec for the reference's Snapshot and
// Probability outputs and for resume (host code; no GPU, no NetCDF library).
//
// Format: the NetCDF classic format, version 2 (64-bit offsets) for writing;
// versions 1 and 2 for reading.  A file is a header (magic "CDF\x02",
// numrecs, dimension list, global attributes, variable list with each
// variable's shape, attributes, type, per-record size and begin offset),
// then the non-record variables, then the records: record i holds record i
// of every record variable in definition order, each padded to 4 bytes.  All
// numbers are big-endian.
//
// Schemas follow src/IO_f.cpp: Snapshot (:90-132) and Probability (:136-160).
#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "xf_internal.h"
#include "../../include/xflat_io.h"

namespace {

constexpr uint32_t NC_DIMENSION = 0x0A, NC_VARIABLE = 0x0B, NC_ATTRIBUTE = 0x0C;
constexpr uint32_t NC_CHAR = 2, NC_INT = 4, NC_DOUBLE = 6;

// ---- big-endian writer -----------------------------------------------------
struct Out {
    std::vector<unsigned char> b;
    void u32(uint32_t v) { for (int s = 24; s >= 0; s -= 8) b.push_back((unsigned char)(v >> s)); }
    void u64(uint64_t v) { for (int s = 56; s >= 0; s -= 8) b.push_back((unsigned char)(v >> s)); }
    void f64(double d) { uint64_t v; std::memcpy(&v, &d, 8); u64(v); }
    void pad4() { while (b.size() & 3) b.push_back(0); }
    void name(const std::string& s) { u32((uint32_t)s.size()); b.insert(b.end(), s.begin(), s.end()); pad4(); }
};

inline void put_be64(unsigned char* p, uint64_t v) {
    for (int i = 7; i >= 0; --i) { p[i] = (unsigned char)v; v >>= 8; }
}
inline uint64_t get_be64(const unsigned char* p) {
    uint64_t v = 0;
    for (int i = 0; i < 8; ++i) v = (v << 8) | p[i];
    return v;
}
inline uint32_t get_be32(const unsigned char* p) {
    return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3];
}

struct Dim { std::string name; uint64_t len; };       // len 0 = record dimension
struct Att { std::string name; std::vector<double> v; };
struct Var {
    std::string name;
    std::vector<int> dims;
    std::vector<Att> atts;
    uint64_t vsize = 0;      // bytes per record (record variable) or total
    uint64_t begin = 0;
    bool rec = false;
};

}  // namespace

struct xf_nc {
    FILE* fp = nullptr;
    std::vector<Dim> dims;
    std::vector<Var> vars;
    uint64_t recsize = 0;      // bytes of one record (all record variables)
    uint64_t data0 = 0;        // offset of record 0
    int64_t nrec = 0;
    int kind = 0;              // 0 snapshot, 1 probability
    int nbeam = 0, ne = 0, nth = 0, nph = 0;
    std::vector<unsigned char> buf;
};

namespace {

int nc_fail_io(const char* what, const char* path) {
    return xf_fail(XF_EIO, "%s %s: %s", what, path ? path : "", std::strerror(errno));
}

// Header of the schema; begin offsets are assigned after the header size is
// known (it depends only on the names, so two passes give the same size).
std::vector<unsigned char> header(const xf_nc& f) {
    Out o;
    o.b.push_back('C'); o.b.push_back('D'); o.b.push_back('F'); o.b.push_back(2);
    o.u32((uint32_t)f.nrec);
    o.u32(NC_DIMENSION);
    o.u32((uint32_t)f.dims.size());
    for (const Dim& d : f.dims) { o.name(d.name); o.u32((uint32_t)d.len); }
    o.u32(0); o.u32(0);                                   // no global attributes
    o.u32(NC_VARIABLE);
    o.u32((uint32_t)f.vars.size());
    for (const Var& v : f.vars) {
        o.name(v.name);
        o.u32((uint32_t)v.dims.size());
        for (int d : v.dims) o.u32((uint32_t)d);
        if (v.atts.empty()) { o.u32(0); o.u32(0); }
        else {
            o.u32(NC_ATTRIBUTE);
            o.u32((uint32_t)v.atts.size());
            for (const Att& a : v.atts) {
                o.name(a.name);
                o.u32(NC_DOUBLE);
                o.u32((uint32_t)a.v.size());
                for (double x : a.v) o.f64(x);
            }
        }
        o.u32(NC_DOUBLE);
        o.u32((uint32_t)std::min<uint64_t>(v.vsize, 0xFFFFFFFFull));
        o.u64(v.begin);
    }
    return o.b;
}

int finish_schema(xf_nc& f, const char* path) {
    // record variables only in these schemas: begins follow the header
    f.recsize = 0;
    for (Var& v : f.vars) {
        uint64_t n = 8;
        for (int d : v.dims) if (f.dims[d].len) n *= f.dims[d].len;
        v.vsize = (n + 3) & ~3ull;
        v.rec = !v.dims.empty() && f.dims[v.dims[0]].len == 0;
    }
    const uint64_t hsize = header(f).size();
    uint64_t off = hsize;
    for (Var& v : f.vars) { v.begin = off; off += v.vsize; f.recsize += v.vsize; }
    f.data0 = hsize;
    f.fp = std::fopen(path, "wb+");
    if (!f.fp) return nc_fail_io("cannot create", path);
    std::vector<unsigned char> h = header(f);
    if (std::fwrite(h.data(), 1, h.size(), f.fp) != h.size()) return nc_fail_io("cannot write", path);
    std::fflush(f.fp);
    return XF_OK;
}

int write_numrecs(xf_nc* f) {
    unsigned char b[4] = {(unsigned char)(f->nrec >> 24), (unsigned char)(f->nrec >> 16), (unsigned char)(f->nrec >> 8),
                          (unsigned char)f->nrec};
    if (std::fseek(f->fp, 4, SEEK_SET) != 0 || std::fwrite(b, 1, 4, f->fp) != 4)
        return xf_fail(XF_EIO, "cannot update the record count");
    return XF_OK;
}

// Writes one record: values of every record variable in order (big-endian).
int put_record(xf_nc* f, const std::vector<const double*>& src) {
    const uint64_t off = f->data0 + (uint64_t)f->nrec * f->recsize;
    if (std::fseek(f->fp, (long)off, SEEK_SET) != 0) return xf_fail(XF_EIO, "seek failed");
    for (size_t i = 0; i < f->vars.size(); ++i) {
        const Var& v = f->vars[i];
        const uint64_t n = v.vsize / 8;
        f->buf.resize(v.vsize);
        for (uint64_t k = 0; k < n; ++k) {
            uint64_t u;
            std::memcpy(&u, src[i] + k, 8);
            put_be64(f->buf.data() + 8 * k, u);
        }
        if (std::fwrite(f->buf.data(), 1, v.vsize, f->fp) != v.vsize) return xf_fail(XF_EIO, "write failed");
    }
    f->nrec += 1;
    return write_numrecs(f);
}

bool check_params(const xf_params* p) {
    if (!p || p->ebins <= 0) return false;
    if (p->module == XF_SA) return true;
    return p->abins > 0 && (p->module == XF_MA || (p->module == XF_MAA && p->pbins > 0));
}
// beams of the state: theta x phi (MA, MAA); one trajectory for SA
int n_theta(const xf_params* p) { return p->module == XF_SA ? 1 : p->abins; }
int n_phi(const xf_params* p) { return p->module == XF_MAA ? p->pbins : 1; }

}  // namespace

extern "C" {

int xf_output_name(const xf_params* p, int probability, int ctr, int rank, char* out, size_t cap) {
    if (!p || !out) return xf_fail(XF_EINVAL, "xf_output_name: bad arguments");
    int n = std::snprintf(out, cap, "%s%s%d_%d_.nc", p->file_prefix, probability ? "Probability" : "Snapshot", ctr, rank);
    if (n < 0 || (size_t)n >= cap) return xf_fail(XF_EINVAL, "xf_output_name: buffer too small");
    return XF_OK;
}

int xf_snapshot_create(const xf_params* p, const char* path, xf_nc** out) {
    if (!check_params(p) || !path || !out) return xf_fail(XF_EINVAL, "xf_snapshot_create: bad arguments");
    std::unique_ptr<xf_nc> f(new xf_nc());
    const bool maa = p->module == XF_MAA;
    f->kind = 0;
    f->nth = n_theta(p);
    f->nph = n_phi(p);
    f->ne = p->ebins;
    f->nbeam = f->nth * f->nph;
    // getDimInfo: Phy_MA.cpp:92-99, Phy_MAA.cpp:94-102, Phy_SA.cpp:74-80 (no theta)
    f->dims.push_back({"r", 0});
    if (p->module != XF_SA) f->dims.push_back({"theta", (uint64_t)f->nth});
    if (maa) f->dims.push_back({"phi", (uint64_t)f->nph});
    f->dims.push_back({"prtcl", (uint64_t)xfc::NPR});
    f->dims.push_back({"comp", 4});
    f->dims.push_back({"ebin", (uint64_t)f->ne});
    Var r{"r", {0}, {}}, dr{"dr", {0}, {}}, psi{"psi", {}, {}};
    for (int d = 0; d < (int)f->dims.size(); ++d) psi.dims.push_back(d);
    // attributes of IO_f.cpp:119-128, names as the reference forms them:
    // string("..." + i) offsets the literal by i characters
    psi.atts.push_back({"E0", {p->E0}});
    psi.atts.push_back({"E1", {p->E1}});
    const char* t_p = "Teprature for particle";
    const char* t_a = "Teprature for anti-particle";
    const char* e_p = "eta for particle";
    const char* e_a = "eta for anti-particle";
    const int flvs = p->flvs > 0 ? p->flvs : 2;
    for (int i = 0; i < flvs; ++i) {
        psi.atts.push_back({std::string(t_p + i), {p->T[i]}});
        psi.atts.push_back({std::string(t_a + i), {p->T[i + 1]}});
        psi.atts.push_back({std::string(e_p + i), {p->eta[i]}});
        psi.atts.push_back({std::string(e_a + i), {p->eta[i + 1]}});
    }
    f->vars = {r, dr, psi};
    int rc = finish_schema(*f, path);
    if (rc) { if (f->fp) std::fclose(f->fp); return rc; }
    *out = f.release();
    return XF_OK;
}

int xf_snapshot_append(xf_nc* f, double r, double dr, const double* psi) {
    if (!f || f->kind != 0 || !psi) return xf_fail(XF_EINVAL, "xf_snapshot_append: bad arguments");
    return put_record(f, {&r, &dr, psi});
}

int xf_probability_create(const xf_params* p, const char* path, xf_nc** out) {
    if (!check_params(p) || !path || !out) return xf_fail(XF_EINVAL, "xf_probability_create: bad arguments");
    std::unique_ptr<xf_nc> f(new xf_nc());
    f->kind = 1;
    // theta = beamLen / (Pbins * 4), phi = Pbins (IO_f.cpp:93-94,164-166): the
    // beams of the state; for SA one trajectory
    f->nth = n_theta(p);
    f->nph = n_phi(p);
    f->ne = p->ebins;
    f->nbeam = f->nth * f->nph;
    f->dims = {{"r", 0}, {"theta", (uint64_t)f->nth}, {"phi", (uint64_t)f->nph}};
    f->vars.push_back({"r", {0}, {}});
    for (int q = 0; q < xfc::NPR; ++q) f->vars.push_back({"prtcl" + std::to_string(q), {0, 1, 2}, {}});
    int rc = finish_schema(*f, path);
    if (rc) { if (f->fp) std::fclose(f->fp); return rc; }
    *out = f.release();
    return XF_OK;
}

int xf_probability_append(xf_nc* f, double r, const double* psi, const double* fv) {
    if (!f || f->kind != 1 || !psi || !fv) return xf_fail(XF_EINVAL, "xf_probability_append: bad arguments");
    const int ne = f->ne;
    std::vector<double> sums((size_t)xfc::NPR * f->nbeam);
    for (int b = 0; b < f->nbeam; ++b)
        for (int q = 0; q < xfc::NPR; ++q) {
            // sum_e |a|^2 fv of species q % 2 (IO_f.cpp:346-351)
            const double* ar = psi + (((size_t)b * xfc::NPR + q) * 4 + 0) * ne;
            const double* ai = ar + ne;
            const double* w = fv + (size_t)(q % 2) * ne;
            double s = 0.;
            for (int e = 0; e < ne; ++e) s += (ar[e] * ar[e] + ai[e] * ai[e]) * w[e];
            sums[(size_t)q * f->nbeam + b] = s;
        }
    std::vector<const double*> src{&r};
    for (int q = 0; q < xfc::NPR; ++q) src.push_back(sums.data() + (size_t)q * f->nbeam);
    return put_record(f, src);
}

int xf_nc_records(const xf_nc* f, int64_t* n) {
    if (!f || !n) return xf_fail(XF_EINVAL, "xf_nc_records: bad arguments");
    *n = f->nrec;
    return XF_OK;
}

int xf_nc_sync(xf_nc* f) {
    if (!f) return xf_fail(XF_EINVAL, "xf_nc_sync: null file");
    int rc = write_numrecs(f);
    if (rc) return rc;
    if (std::fflush(f->fp) != 0) return xf_fail(XF_EIO, "flush failed");
    return XF_OK;
}

int xf_nc_close(xf_nc* f) {
    if (!f) return XF_OK;
    int rc = f->fp ? xf_nc_sync(f) : XF_OK;
    if (f->fp && std::fclose(f->fp) != 0 && rc == XF_OK) rc = xf_fail(XF_EIO, "close failed");
    delete f;
    return rc;
}

// ---- reader ------------------------------------------------------------------
int xf_snapshot_read_last(const char* path, const xf_params* p, double* psi, double* r, double* dr) {
    if (!path || !check_params(p) || !psi) return xf_fail(XF_EINVAL, "xf_snapshot_read_last: bad arguments");
    FILE* fp = std::fopen(path, "rb");
    if (!fp) return nc_fail_io("cannot open", path);
    std::vector<unsigned char> b;
    {
        unsigned char chunk[1 << 16];
        // the header is at most a few KB for these schemas; read 1 MB at most
        size_t n;
        while (b.size() < (1u << 20) && (n = std::fread(chunk, 1, sizeof chunk, fp)) > 0) b.insert(b.end(), chunk, chunk + n);
    }
    auto bad = [&](const char* what) { std::fclose(fp); return xf_fail(XF_EIO, "%s: %s", path, what); };
    size_t pos = 0;
    auto need = [&](size_t n) { return pos + n <= b.size(); };
    auto u32 = [&]() { uint32_t v = get_be32(&b[pos]); pos += 4; return v; };
    if (!need(8) || b[0] != 'C' || b[1] != 'D' || b[2] != 'F' || (b[3] != 1 && b[3] != 2)) return bad("not a NetCDF classic file");
    const int ver = b[3];
    pos = 4;
    uint32_t numrecs = u32();
    auto read_name = [&](std::string& s) -> bool {
        if (!need(4)) return false;
        uint32_t n = u32();
        size_t pn = (n + 3) & ~3u;
        if (!need(pn)) return false;
        s.assign((const char*)&b[pos], n);
        pos += pn;
        return true;
    };
    auto skip_atts = [&]() -> bool {
        if (!need(8)) return false;
        uint32_t tag = u32(), n = u32();
        if (tag == 0) return n == 0;
        if (tag != NC_ATTRIBUTE) return false;
        for (uint32_t i = 0; i < n; ++i) {
            std::string s;
            if (!read_name(s) || !need(8)) return false;
            uint32_t t = u32(), cnt = u32();
            size_t es = (t == 1 || t == 2) ? 1 : (t == 3) ? 2 : (t == 4 || t == 5) ? 4 : 8;
            size_t bytes = ((size_t)cnt * es + 3) & ~(size_t)3;
            if (!need(bytes)) return false;
            pos += bytes;
        }
        return true;
    };
    std::vector<Dim> dims;
    if (!need(8)) return bad("truncated header");
    {
        uint32_t tag = u32(), n = u32();
        if (tag != NC_DIMENSION && !(tag == 0 && n == 0)) return bad("bad dimension list");
        for (uint32_t i = 0; i < n; ++i) {
            Dim d;
            if (!read_name(d.name) || !need(4)) return bad("truncated dimension");
            d.len = u32();
            dims.push_back(d);
        }
    }
    if (!skip_atts()) return bad("bad global attributes");
    std::vector<Var> vars;
    std::vector<uint32_t> types;
    if (!need(8)) return bad("truncated header");
    {
        uint32_t tag = u32(), n = u32();
        if (tag != NC_VARIABLE && !(tag == 0 && n == 0)) return bad("bad variable list");
        for (uint32_t i = 0; i < n; ++i) {
            Var v;
            if (!read_name(v.name) || !need(4)) return bad("truncated variable");
            uint32_t nd = u32();
            if (!need(4ull * nd)) return bad("truncated variable");
            for (uint32_t k = 0; k < nd; ++k) v.dims.push_back((int)u32());
            if (!skip_atts() || !need(8)) return bad("bad variable attributes");
            types.push_back(u32());
            v.vsize = u32();
            if (ver == 2) { if (!need(8)) return bad("truncated"); v.begin = get_be64(&b[pos]); pos += 8; }
            else { if (!need(4)) return bad("truncated"); v.begin = u32(); }
            for (int d : v.dims) if (d < 0 || d >= (int)dims.size()) return bad("bad dimension id");
            v.rec = !v.dims.empty() && dims[v.dims[0]].len == 0;
            vars.push_back(v);
        }
    }
    // record size: sum of record variables' sizes (one record variable: unpadded)
    uint64_t recsize = 0;
    int nrecvars = 0;
    for (size_t i = 0; i < vars.size(); ++i) {
        if (!vars[i].rec) continue;
        ++nrecvars;
        uint64_t n = 1;
        size_t es = (types[i] == 1 || types[i] == 2) ? 1 : (types[i] == 3) ? 2 : (types[i] == 4 || types[i] == 5) ? 4 : 8;
        for (int d : vars[i].dims) if (dims[d].len) n *= dims[d].len;
        uint64_t sz = n * es;
        recsize += (sz + 3) & ~3ull;
        vars[i].vsize = sz;
    }
    if (nrecvars == 1) for (size_t i = 0; i < vars.size(); ++i) if (vars[i].rec) recsize = vars[i].vsize;
    auto find = [&](const char* name) -> int {
        for (size_t i = 0; i < vars.size(); ++i) if (vars[i].name == name) return (int)i;
        return -1;
    };
    const int ir = find("r"), idr = find("dr"), ip = find("psi");
    if (ir < 0 || idr < 0 || ip < 0) return bad("no r, dr or psi variable");
    if (types[ir] != NC_DOUBLE || types[idr] != NC_DOUBLE || types[ip] != NC_DOUBLE) return bad("variables are not double");
    if (!vars[ir].rec || !vars[idr].rec || !vars[ip].rec) return bad("r, dr, psi must be record variables");
    if (numrecs == 0 || numrecs == 0xFFFFFFFFu) {
        // streaming or empty: count whole records from the file size
        if (std::fseek(fp, 0, SEEK_END) != 0) return bad("seek failed");
        long sz = std::ftell(fp);
        uint64_t first = vars[ir].begin;
        for (const Var& v : vars) if (v.rec) first = std::min<uint64_t>(first, v.begin);
        numrecs = (sz > (long)first && recsize) ? (uint32_t)(((uint64_t)sz - first) / recsize) : 0;
        if (numrecs == 0) return bad("no records");
    }
    const uint64_t want = (uint64_t)n_theta(p) * n_phi(p) * xfc::NPR * 4 * p->ebins;
    uint64_t have = 1;
    for (size_t k = 1; k < vars[ip].dims.size(); ++k) have *= dims[vars[ip].dims[k]].len;
    if (have != want) return bad("psi shape does not match the configuration");
    const uint64_t last = numrecs - 1;
    auto read_doubles = [&](const Var& v, double* dst, uint64_t n) -> bool {
        std::vector<unsigned char> tmp(n * 8);
        if (std::fseek(fp, (long)(v.begin + last * recsize), SEEK_SET) != 0) return false;
        if (std::fread(tmp.data(), 1, tmp.size(), fp) != tmp.size()) return false;
        for (uint64_t k = 0; k < n; ++k) {
            uint64_t u = get_be64(&tmp[8 * k]);
            std::memcpy(dst + k, &u, 8);
        }
        return true;
    };
    double rr = 0., ddr = 0.;
    if (!read_doubles(vars[ir], &rr, 1) || !read_doubles(vars[idr], &ddr, 1) || !read_doubles(vars[ip], psi, want))
        return bad("truncated record");
    std::fclose(fp);
    if (r) *r = rr;
    if (dr) *dr = ddr;
    return XF_OK;
}

}  // extern "C"
