// Cone-aware permutations of an assembled problem and partition cuts: the
// integer preparation the reference runs before splitting a problem over
// several devices (reference src/permute.cu, src/partition_utils.c,
// distributed/distributed_utils.cu:179-216).  Host code; the random
// permutations draw from the C library's rand() in the reference's order, so
// a caller that seeds with srand() gets the reference's permutation.
#include <algorithm>
#include <cfloat>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <vector>

#include "../../include/pdhcg_partition.h"
#include "host_common.h"

namespace pdhcg {
namespace {

template <class T> T *copy_array(const T *src, size_t count) {
    if (!src) return nullptr;
    T *d = static_cast<T *>(std::malloc(count ? count * sizeof(T) : 1));
    if (!d) throw std::bad_alloc();
    if (count) std::memcpy(d, src, count * sizeof(T));
    return d;
}

CsrComponent *copy_csr(const CsrComponent *c, int rows, int nnz) {
    if (!c) return nullptr;
    auto *d = static_cast<CsrComponent *>(std::calloc(1, sizeof(CsrComponent)));
    d->row_ptr = copy_array(c->row_ptr, (size_t)rows + 1);
    d->col_ind = nnz > 0 ? copy_array(c->col_ind, (size_t)nnz) : nullptr;
    d->val = nnz > 0 ? copy_array(c->val, (size_t)nnz) : nullptr;
    return d;
}

void copy_blocks(cone_blocks_t &d, const cone_blocks_t &s) {
    std::memset(&d, 0, sizeof(d));
    d.num_cones = s.num_cones;
    if (s.num_cones > 0) {
        const size_t k = (size_t)s.num_cones;
        d.start_idx = copy_array(s.start_idx, k);
        d.v_dim = copy_array(s.v_dim, k);
        d.type = copy_array(s.type, k);
        d.power_alpha = copy_array(s.power_alpha, k);
    }
    d.fixed_mask_size = s.fixed_mask_size;
    d.is_fixed = s.fixed_mask_size > 0 ? copy_array(s.is_fixed, (size_t)s.fixed_mask_size) : nullptr;
}

qp_problem_t *copy_problem(const qp_problem_t *p) {
    const int n = p->num_variables, m = p->num_constraints, k = p->num_rank_lowrank_obj;
    auto *d = static_cast<qp_problem_t *>(std::calloc(1, sizeof(qp_problem_t)));
    *d = *p;
    d->variable_lower_bound = copy_array(p->variable_lower_bound, n);
    d->variable_upper_bound = copy_array(p->variable_upper_bound, n);
    d->objective_vector = copy_array(p->objective_vector, n);
    d->constraint_lower_bound = copy_array(p->constraint_lower_bound, m);
    d->constraint_upper_bound = copy_array(p->constraint_upper_bound, m);
    d->affine_cone_offset = copy_array(p->affine_cone_offset, m);
    d->primal_start = copy_array(p->primal_start, n);
    d->dual_start = copy_array(p->dual_start, m);
    d->constraint_matrix = copy_csr(p->constraint_matrix, m, p->constraint_matrix_num_nonzeros);
    d->objective_sparse_matrix = copy_csr(p->objective_sparse_matrix, n, p->objective_sparse_matrix_num_nonzeros);
    d->objective_lowrank_matrix = copy_csr(p->objective_lowrank_matrix, k, p->objective_lowrank_matrix_num_nonzeros);
    d->objective_lowrank_middle_matrix =
        copy_csr(p->objective_lowrank_middle_matrix, k, p->objective_lowrank_middle_matrix_num_nonzeros);
    copy_blocks(d->cones, p->cones);
    copy_blocks(d->affine_cones, p->affine_cones);
    d->num_quadratic_constraints = 0;
    d->quadratic_constraint_row_indices = nullptr;
    d->quadratic_constraint_matrices = nullptr;
    d->quadratic_constraint_matrix_num_nonzeros = nullptr;
    return d;
}

// inverse[perm[i]] = i; empty when perm is not a permutation of 0..n-1.
bool inverse_of(int n, const int *perm, std::vector<int> &inv) {
    inv.assign((size_t)std::max(n, 0), -1);
    if (n <= 0) return true;
    if (!perm) return false;
    for (int i = 0; i < n; ++i) {
        const int v = perm[i];
        if (v < 0 || v >= n || inv[v] >= 0) return false;
        inv[v] = i;
    }
    return true;
}

// True when every block lands on consecutive positions in its own order.
bool keeps_blocks(const cone_blocks_t &b, int n, const int *perm) {
    std::vector<int> inv;
    if (!inverse_of(n, perm, inv)) return false;
    for (int k = 0; k < b.num_cones; ++k) {
        const int s = b.start_idx[k], len = cone_block_slots(&b, k), to = inv[s];
        for (int j = 1; j < len; ++j)
            if (inv[s + j] != to + j) return false;
    }
    return true;
}

template <class T> void gather_in_place(T *a, int n, const int *perm) {
    if (!a || n <= 0) return;
    std::vector<T> t((size_t)n);
    for (int i = 0; i < n; ++i) t[i] = a[perm[i]];
    std::copy(t.begin(), t.end(), a);
}

// New row i is old row perm[i].
void reorder_rows(CsrComponent *c, int rows, int nnz, const int *perm) {
    if (!c || nnz == 0) return;
    std::vector<int> rp((size_t)rows + 1), ci((size_t)nnz);
    std::vector<double> v((size_t)nnz);
    int at = 0;
    for (int i = 0; i < rows; ++i) {
        const int s = c->row_ptr[perm[i]], e = c->row_ptr[perm[i] + 1];
        std::copy(c->col_ind + s, c->col_ind + e, ci.begin() + at);
        std::copy(c->val + s, c->val + e, v.begin() + at);
        at += e - s;
        rp[(size_t)i + 1] = at;
    }
    std::copy(rp.begin(), rp.end(), c->row_ptr);
    std::copy(ci.begin(), ci.end(), c->col_ind);
    std::copy(v.begin(), v.end(), c->val);
}

// Columns renamed through old_to_new, each row re-sorted by column (the
// columns of a row are distinct, so the order is unique).
void rename_columns(int rows, const int *rp, int *ci, double *v, const int *old_to_new) {
    std::vector<std::pair<int, double>> row;
    for (int r = 0; r < rows; ++r) {
        const int s = rp[r], e = rp[r + 1];
        row.clear();
        for (int j = s; j < e; ++j) row.emplace_back(old_to_new[ci[j]], v[j]);
        std::sort(row.begin(), row.end(), [](const auto &a, const auto &b) { return a.first < b.first; });
        for (int j = s; j < e; ++j) {
            ci[j] = row[(size_t)(j - s)].first;
            v[j] = row[(size_t)(j - s)].second;
        }
    }
}

struct Unit {
    int start, length;
};

// Fisher-Yates with rand() % (i + 1) from the last position down.
template <class T> void shuffle_with_rand(std::vector<T> &a) {
    for (int i = (int)a.size() - 1; i > 0; --i) std::swap(a[(size_t)i], a[(size_t)(std::rand() % (i + 1))]);
}

void vector_permutation(int n, const cone_blocks_t &b, int method, int block_size, int *perm) {
    if (method == NO_PERMUTATION || n <= 1) {
        std::iota(perm, perm + std::max(n, 0), 0);
        return;
    }
    const int free_len = method == FULL_RANDOM_PERMUTATION ? 1 : std::max(block_size, 1);
    std::vector<Unit> units;
    if (b.num_cones <= 0) {
        if (method == FULL_RANDOM_PERMUTATION) {
            std::vector<int> p((size_t)n);
            std::iota(p.begin(), p.end(), 0);
            shuffle_with_rand(p);
            std::copy(p.begin(), p.end(), perm);
            return;
        }
        std::vector<int> ids((size_t)((n + free_len - 1) / free_len));
        std::iota(ids.begin(), ids.end(), 0);
        shuffle_with_rand(ids);
        int at = 0;
        for (int id : ids)
            for (int v = id * free_len; v < std::min((id + 1) * free_len, n); ++v) perm[at++] = v;
        return;
    }
    std::vector<Unit> cones((size_t)b.num_cones);
    for (int k = 0; k < b.num_cones; ++k) cones[k] = {b.start_idx[k], cone_block_slots(&b, k)};
    std::sort(cones.begin(), cones.end(), [](const Unit &a, const Unit &c) { return a.start < c.start; });
    int cursor = 0;
    auto free_run = [&](int until) {
        while (cursor < until) {
            const int len = std::min(free_len, until - cursor);
            units.push_back({cursor, len});
            cursor += len;
        }
    };
    for (const Unit &c : cones) {
        free_run(c.start);
        units.push_back(c);
        cursor = c.start + c.length;
    }
    free_run(n);
    shuffle_with_rand(units);
    int at = 0;
    for (const Unit &u : units)
        for (int j = 0; j < u.length; ++j) perm[at++] = u.start + j;
}

// Legal cut positions are 1..total-1 outside the merged forbidden spans.
struct Boundaries {
    int total;
    std::vector<std::pair<int, int>> spans;  // sorted, merged [first, last]

    int nearest(int at, int dir) const {
        if (total <= 1) return -1;
        if (at < 1) {
            if (dir < 0) return -1;
            at = 1;
        }
        if (at >= total) {
            if (dir > 0) return -1;
            at = total - 1;
        }
        auto it = std::upper_bound(spans.begin(), spans.end(), at,
                                   [](int x, const std::pair<int, int> &s) { return x < s.first; });
        if (it != spans.begin()) {
            const auto &s = *(it - 1);
            if (at <= s.second) at = dir < 0 ? s.first - 1 : s.second + 1;
        }
        return (at > 0 && at < total) ? at : -1;
    }
};

}  // namespace
}  // namespace pdhcg

using namespace pdhcg;

extern "C" {

bool validate_cone_permutation(const qp_problem_t *qp, const int *col_perm) {
    return qp && keeps_blocks(qp->cones, qp->num_variables, col_perm);
}

bool validate_affine_cone_row_permutation(const qp_problem_t *qp, const int *row_perm) {
    return qp && keeps_blocks(qp->affine_cones, qp->num_constraints, row_perm);
}

bool permute_problem(qp_problem_t *qp, int *row_perm, int *col_perm) {
    if (!qp || (qp->num_constraints > 0 && !row_perm) || (qp->num_variables > 0 && !col_perm)) return false;
    const int m = qp->num_constraints, n = qp->num_variables;
    if (!validate_cone_permutation(qp, col_perm)) {
        std::fprintf(stderr, "Error: column permutation splits or reorders a cone block.\n");
        return false;
    }
    if (!validate_affine_cone_row_permutation(qp, row_perm)) {
        std::fprintf(stderr, "Error: row permutation splits or reorders an affine cone block.\n");
        return false;
    }
    for (double *a : {qp->objective_vector, qp->variable_lower_bound, qp->variable_upper_bound, qp->primal_start})
        gather_in_place(a, n, col_perm);
    for (double *a : {qp->constraint_lower_bound, qp->constraint_upper_bound, qp->affine_cone_offset, qp->dual_start})
        gather_in_place(a, m, row_perm);
    std::vector<int> inv_row, inv_col;
    inverse_of(m, row_perm, inv_row);
    inverse_of(n, col_perm, inv_col);
    for (int k = 0; k < qp->affine_cones.num_cones; ++k)
        qp->affine_cones.start_idx[k] = inv_row[qp->affine_cones.start_idx[k]];
    if (qp->constraint_matrix && qp->constraint_matrix_num_nonzeros > 0) {
        reorder_rows(qp->constraint_matrix, m, qp->constraint_matrix_num_nonzeros, row_perm);
        rename_columns(m, qp->constraint_matrix->row_ptr, qp->constraint_matrix->col_ind, qp->constraint_matrix->val,
                       inv_col.data());
    }
    if (qp->objective_sparse_matrix && qp->objective_sparse_matrix_num_nonzeros > 0) {
        reorder_rows(qp->objective_sparse_matrix, n, qp->objective_sparse_matrix_num_nonzeros, col_perm);
        rename_columns(n, qp->objective_sparse_matrix->row_ptr, qp->objective_sparse_matrix->col_ind,
                       qp->objective_sparse_matrix->val, inv_col.data());
    }
    if (qp->objective_lowrank_matrix && qp->objective_lowrank_matrix_num_nonzeros > 0)
        rename_columns(qp->num_rank_lowrank_obj, qp->objective_lowrank_matrix->row_ptr,
                       qp->objective_lowrank_matrix->col_ind, qp->objective_lowrank_matrix->val, inv_col.data());
    for (int k = 0; k < qp->cones.num_cones; ++k) qp->cones.start_idx[k] = inv_col[qp->cones.start_idx[k]];
    if (qp->cones.num_cones > 0 && qp->cones.is_fixed) gather_in_place(qp->cones.is_fixed, n, col_perm);
    return true;
}

qp_problem_t *permute_problem_return_new(const qp_problem_t *qp, int *row_perm, int *col_perm) {
    if (!qp) return nullptr;
    qp_problem_t *d = nullptr;
    try {
        d = copy_problem(qp);
    } catch (const std::exception &) {
        qp_problem_free(d);
        return nullptr;
    }
    if (!permute_problem(d, row_perm, col_perm)) {
        qp_problem_free(d);
        return nullptr;
    }
    return d;
}

void generate_cone_aware_permutation(const qp_problem_t *qp, permute_method_t method, int block_size, int *perm) {
    vector_permutation(qp->num_variables, qp->cones, (int)method, block_size, perm);
}

void generate_affine_cone_aware_row_permutation(const qp_problem_t *qp, permute_method_t method, int block_size,
                                                int *perm) {
    vector_permutation(qp->num_constraints, qp->affine_cones, (int)method, block_size, perm);
}

void repermute_solution(pdhcg_result_t *result, int *row_perm, int *col_perm) {
    std::vector<int> inv_row, inv_col;
    inverse_of(result->num_variables, col_perm, inv_col);
    inverse_of(result->num_constraints, row_perm, inv_row);
    gather_in_place(result->primal_solution, result->num_variables, inv_col.data());
    gather_in_place(result->reduced_cost, result->num_variables, inv_col.data());
    gather_in_place(result->dual_solution, result->num_constraints, inv_row.data());
}

int *get_balanced_cuts(const int *weights, int total_dim, int num_partitions) {
    int *cuts = static_cast<int *>(std::malloc(((size_t)num_partitions + 1) * sizeof(int)));
    if (!cuts) return nullptr;
    cuts[0] = 0;
    cuts[num_partitions] = total_dim;
    if (num_partitions == 1) return cuts;
    long long total = 0;
    for (int i = 0; i < total_dim; ++i) total += weights[i];
    const double per_part = (double)total / num_partitions;
    long long run = 0;
    int part = 1;
    for (int i = 0; i < total_dim && part < num_partitions; ++i) {
        run += weights[i];
        if (run >= part * per_part) cuts[part++] = i + 1;
    }
    for (; part < num_partitions; ++part) cuts[part] = total_dim;
    return cuts;
}

bool optimize_partition_cuts(int total_dimension, int num_partitions, const int *forbidden_starts,
                             const int *forbidden_ends, int num_forbidden_intervals, int *cuts) {
    if (total_dimension < 0 || num_partitions <= 0 || num_forbidden_intervals < 0 || !cuts ||
        (num_forbidden_intervals > 0 && (!forbidden_starts || !forbidden_ends)))
        return false;
    const int need = num_partitions - 1;
    if (need <= 0) return true;
    if (total_dimension < num_partitions) return false;
    Boundaries B{total_dimension, {}};
    std::vector<std::pair<int, int>> raw;
    for (int i = 0; i < num_forbidden_intervals; ++i) {
        const int f = forbidden_starts[i], l = forbidden_ends[i];
        if (f < 1 || l < f || l >= total_dimension) return false;
        raw.emplace_back(f, l);
    }
    std::sort(raw.begin(), raw.end());
    for (const auto &s : raw) {
        if (B.spans.empty() || s.first > B.spans.back().second + 1) B.spans.push_back(s);
        else B.spans.back().second = std::max(B.spans.back().second, s.second);
    }
    // candidates: up to P legal boundaries on each side of every target and
    // of both ends (an optimum cannot skip more)
    std::vector<int> cand;
    auto walk = [&](int from, int dir) {
        int at = B.nearest(from, dir);
        for (int s = 0; s < num_partitions && at > 0; ++s) {
            cand.push_back(at);
            at = B.nearest(at + dir, dir);
        }
    };
    for (int c = 1; c < num_partitions; ++c) {
        walk(cuts[c], -1);
        walk(cuts[c], 1);
    }
    walk(1, 1);
    walk(total_dimension - 1, -1);
    std::sort(cand.begin(), cand.end());
    cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
    const int nc = (int)cand.size();
    if (nc < need) return false;
    // cost[c][j]: best squared displacement of cuts 1..c+1 with cut c+1 at cand[j]
    std::vector<long double> cost((size_t)need * nc);
    std::vector<int> from((size_t)need * nc, -1);
    for (int j = 0; j < nc; ++j) {
        const long double d = (long double)cand[j] - cuts[1];
        cost[j] = d * d;
    }
    for (int c = 1; c < need; ++c) {
        long double best = LDBL_MAX;
        int arg = -1;
        for (int j = 0; j < nc; ++j) {
            if (j > 0 && cost[(size_t)(c - 1) * nc + j - 1] < best) {
                best = cost[(size_t)(c - 1) * nc + j - 1];
                arg = j - 1;
            }
            const size_t e = (size_t)c * nc + j;
            if (arg < 0) {
                cost[e] = LDBL_MAX;
                continue;
            }
            const long double d = (long double)cand[j] - cuts[c + 1];
            cost[e] = best + d * d;
            from[e] = arg;
        }
    }
    long double best = LDBL_MAX;
    int arg = -1;
    for (int j = 0; j < nc; ++j)
        if (cost[(size_t)(need - 1) * nc + j] < best) {
            best = cost[(size_t)(need - 1) * nc + j];
            arg = j;
        }
    for (int c = need - 1; c >= 0 && arg >= 0; --c) {
        cuts[c + 1] = cand[arg];
        arg = from[(size_t)c * nc + arg];
    }
    return best < LDBL_MAX;
}

}  // extern "C"
