/* Reader of free-format MPS / QPS files (ROWS, COLUMNS, RHS, RANGES, BOUNDS,
 * OBJSENSE, QUADOBJ, QMATRIX, QCMATRIX), plain or gzip-compressed.  Same entry
 * point as the reference's include/mps_parser.h (src/mps_parser.c:420).
 * Returns an assembled problem (free with qp_problem_free), or NULL with a
 * message on stderr. */
#pragma once

#include "pdhcg_types.h"

#ifdef __cplusplus
extern "C" {
#endif

qp_problem_t *read_mps_file(const char *filename);

#ifdef __cplusplus
}
#endif
