/*
 * xflat_io.h -- host-side output and resume for the evolution step: the
 * reference's NetCDF "Snapshot" and "Probability" files (src/IO_f.cpp:90-194,
 * :266-384) and resume from a Snapshot (src/IO_f.cpp:197-250), written and
 * read by the project's own NetCDF classic-format (CDF-2, 64-bit offset)
 * codec, so no NetCDF library is needed.  Files are readable by any NetCDF
 * reader; the schema (dimension, variable and attribute names, shapes,
 * record layout) is the reference's.
 *
 * All arrays use the reference's snapshot layout psi[beam][prtcl][comp][ebin]
 * with beam = theta*Pbins + phi (see xflat.h).  Status codes are xflat.h's.
 */
#ifndef XFLAT_IO_H_
#define XFLAT_IO_H_

#include "xflat.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct xf_nc xf_nc;

/* Snapshot file (dumpMode bit 0; io_f::init, IO_f.cpp:90-132): dims r
 * (unlimited), theta, [phi,] prtcl, comp, ebin; variables r[r], dr[r],
 * psi[r][theta][phi][prtcl][comp][ebin] with the E0/E1/temperature/eta
 * attributes of IO_f.cpp:119-128 (including their names). */
int xf_snapshot_create(const xf_params* p, const char* path, xf_nc** out);
/* Appends one record (io_f::writeToFile, IO_f.cpp:266-328). */
int xf_snapshot_append(xf_nc* f, double r, double dr, const double* psi);

/* Probability file (dumpMode bit 1; IO_f.cpp:136-160): dims r (unlimited),
 * theta, phi; variables r[r] and prtcl0..prtcl3[r][theta][phi] =
 * sum_e |a|^2 fv of species (p % 2) (IO_f.cpp:331-384).  fv: [4][ebins]. */
int xf_probability_create(const xf_params* p, const char* path, xf_nc** out);
int xf_probability_append(xf_nc* f, double r, const double* psi, const double* fv);

int xf_nc_records(const xf_nc* f, int64_t* n);
int xf_nc_sync(xf_nc* f);                 /* nc_sync: header numrecs + flush    */
int xf_nc_close(xf_nc* f);

/* Resume (io_f::fillInitData, IO_f.cpp:197-250): reads the last record of the
 * Snapshot file's r, dr and psi.  Accepts CDF-1 and CDF-2 files; checks that
 * psi's shape matches p. */
int xf_snapshot_read_last(const char* path, const xf_params* p, double* psi, double* r, double* dr);

/* File name of the reference: <prefix>Snapshot<ctr>_<rank>_.nc /
 * <prefix>Probability<ctr>_<rank>_.nc (IO_f.cpp:99, :140). */
int xf_output_name(const xf_params* p, int probability, int ctr, int rank, char* out, size_t cap);

#ifdef __cplusplus
}
#endif

#endif /* XFLAT_IO_H_ */
