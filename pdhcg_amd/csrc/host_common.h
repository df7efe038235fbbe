// Host-side helpers shared by the C++ sources of libpdhcg_amd.so.
#pragma once

#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>

#include "../../include/pdhcg.h"

namespace pdhcg {

// Library-internal failure; converted to the C interface's NULL / nonzero
// convention at the exported entry points (the process is never terminated).
struct Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

[[noreturn]] inline void fail(const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    throw Error(buf);
}

// malloc that reports exhaustion as an Error; result owned by C callers (free()).
template <class T> T *cmalloc(size_t count) {
    void *p = std::malloc(count ? count * sizeof(T) : 1);
    if (!p) fail("host allocation of %zu bytes failed", count * sizeof(T));
    return static_cast<T *>(p);
}
template <class T> T *ccalloc(size_t count) {
    void *p = std::calloc(count ? count : 1, sizeof(T));
    if (!p) fail("host allocation of %zu bytes failed", count * sizeof(T));
    return static_cast<T *>(p);
}

// Length in slots of a cone block (reference src/cone_utils.c:26-38).
int cone_slots(cone_type_t type, int v_dim);
int cone_block_slots(const cone_blocks_t *blocks, int k);

// Defaults (reference src/utils.cu:324-368).
void fill_default_parameters(pdhg_parameters_t *p);

// Fixed-section validation (reference src/pdhcg.c:729-892).
int validate_fixed_sections(const qp_problem_t *problem);

// Diagonal quadratic constraints lifted into SOC / RSOC blocks, and the
// result of the lifted problem cut back to the original dimensions
// (reference src/qcqp_transform.c; host_qcqp.cpp).  NULL with a message when
// the problem cannot be lifted.
qp_problem_t *lift_quadratic_constraints(const qp_problem_t *problem, cone_type_t form);
void restore_result_dimensions(pdhcg_result_t *result, const qp_problem_t *original);

// Cancellation request set by pdhcg_request_cancel (reference src/pdhcg.c:31);
// defined in solver.hip, read by every solve loop between evaluations.
extern volatile int g_pdhcg_cancel;

}  // namespace pdhcg
