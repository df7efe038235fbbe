/* Reader of CBF files (VER, OBJSENSE, VAR, CON, OBJACOORD, OBJBCOORD,
 * OBJQCOORD, ACOORD, BCOORD; cones F, L=, L+, L-, Q, QR, EXP), plain or
 * gzip-compressed.  Same entry point as the reference's include/cbf_parser.h
 * (src/cbf_parser.c:1178).  Returns an assembled problem (free with
 * qp_problem_free), or NULL with a message on stderr. */
#pragma once

#include "pdhcg_types.h"

#ifdef __cplusplus
extern "C" {
#endif

qp_problem_t *read_cbf_file(const char *filename);

#ifdef __cplusplus
}
#endif
