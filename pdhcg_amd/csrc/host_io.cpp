// Problem-file readers: free-format MPS / QPS (with QCMATRIX quadratic
// constraints) and CBF, plain or gzip-compressed (zlib reads both).  Entry
// points and results follow the reference's include/mps_parser.h and
// include/cbf_parser.h (src/mps_parser.c, src/cbf_parser.c): the same
// sections, defaults, orderings and rejections, so a file produces the same
// qp_problem_t.
#include <zlib.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "host_common.h"

namespace pdhcg {
namespace {

// Lines of a text file of any length; gzopen reads uncompressed files too.
class LineReader {
  public:
    explicit LineReader(const char *path) : f_(gzopen(path, "rb")) {}
    ~LineReader() {
        if (f_) gzclose(f_);
    }
    bool ok() const { return f_ != nullptr; }
    bool next(std::string &line) {
        line.clear();
        char buf[65536];
        while (gzgets(f_, buf, sizeof(buf))) {
            line += buf;
            if (!line.empty() && line.back() == '\n') return true;
        }
        return !line.empty();
    }

  private:
    gzFile f_;
};

struct Coo {
    std::vector<int> r, c;
    std::vector<double> v;
    void add(int i, int j, double x) {
        r.push_back(i);
        c.push_back(j);
        v.push_back(x);
    }
    size_t nnz() const { return v.size(); }
};

// CSR in input order within each row (a stable counting sort by row).
CsrComponent *coo_rows_stable(const Coo &m, int rows) {
    CsrComponent *c = ccalloc<CsrComponent>(1);
    c->row_ptr = ccalloc<int>((size_t)rows + 1);
    for (int r : m.r) c->row_ptr[r + 1]++;
    for (int i = 0; i < rows; ++i) c->row_ptr[i + 1] += c->row_ptr[i];
    if (m.nnz()) {
        c->col_ind = cmalloc<int>(m.nnz());
        c->val = cmalloc<double>(m.nnz());
        std::vector<int> pos(c->row_ptr, c->row_ptr + rows);
        for (size_t k = 0; k < m.nnz(); ++k) {
            const int d = pos[m.r[k]]++;
            c->col_ind[d] = m.c[k];
            c->val[d] = m.v[k];
        }
    }
    return c;
}

// CSR with each row sorted by column and duplicates summed; zero sums dropped.
CsrComponent *coo_rows_merged(const Coo &m, int rows, int *nnz_out) {
    std::vector<std::vector<std::pair<int, double>>> byrow((size_t)rows);
    for (size_t k = 0; k < m.nnz(); ++k) byrow[m.r[k]].push_back({m.c[k], m.v[k]});
    CsrComponent *c = ccalloc<CsrComponent>(1);
    c->row_ptr = ccalloc<int>((size_t)rows + 1);
    std::vector<int> ci;
    std::vector<double> cv;
    for (int r = 0; r < rows; ++r) {
        auto &e = byrow[r];
        std::stable_sort(e.begin(), e.end(), [](const auto &a, const auto &b) { return a.first < b.first; });
        for (size_t i = 0; i < e.size();) {
            size_t j = i;
            double acc = 0.0;
            while (j < e.size() && e[j].first == e[i].first) acc += e[j++].second;
            if (acc != 0.0) {
                ci.push_back(e[i].first);
                cv.push_back(acc);
            }
            i = j;
        }
        c->row_ptr[r + 1] = (int)ci.size();
    }
    c->col_ind = cmalloc<int>(std::max<size_t>(ci.size(), 1));
    c->val = cmalloc<double>(std::max<size_t>(cv.size(), 1));
    std::copy(ci.begin(), ci.end(), c->col_ind);
    std::copy(cv.begin(), cv.end(), c->val);
    *nnz_out = (int)ci.size();
    return c;
}

template <class T> T *copy_out(const std::vector<T> &v) {
    T *p = cmalloc<T>(std::max<size_t>(v.size(), 1));
    std::copy(v.begin(), v.end(), p);
    return p;
}

// ---------------------------------------------------------------- MPS / QPS

struct Mps {
    enum Sec { NONE, ROWS, COLUMNS, RHS, RANGES, BOUNDS, OBJSENSE, QUADOBJ, QMATRIX, QCMATRIX, ENDATA };
    std::vector<std::pair<char, std::string>> raw_rows;  // until the ROWS section ends
    bool rows_done = false;
    std::string obj_row;
    bool has_obj = false;
    std::unordered_map<std::string, int> rows, cols;
    std::vector<char> row_type;
    std::vector<double> cl, cu, c, lb, ub;
    Coo A, Q;
    struct Qc {
        int row;
        Coo m;
    };
    std::vector<Qc> qcs;
    int qc_cur = -1;
    bool qc_obj = false;
    std::string cur_col;
    bool have_col = false;
    bool maximize = false;
    double obj_const = 0.0;

    int row_of(const std::string &s) const {
        auto it = rows.find(s);
        return it == rows.end() ? -1 : it->second;
    }
    int col_of(const std::string &s) const {
        auto it = cols.find(s);
        return it == cols.end() ? -1 : it->second;
    }
    // the first N row is the objective (the first row at all when there is
    // none); the other E/L/G rows become constraints in order
    void finish_rows() {
        if (rows_done) return;
        rows_done = true;
        int obj = -1;
        for (size_t i = 0; i < raw_rows.size() && obj < 0; ++i)
            if (raw_rows[i].first == 'N') obj = (int)i;
        if (obj < 0 && !raw_rows.empty()) obj = 0;
        if (obj >= 0) {
            obj_row = raw_rows[obj].second;
            has_obj = true;
        }
        for (size_t i = 0; i < raw_rows.size(); ++i) {
            const char t = raw_rows[i].first;
            if ((int)i == obj || (t != 'E' && t != 'L' && t != 'G')) continue;
            if (rows.count(raw_rows[i].second)) continue;
            rows.emplace(raw_rows[i].second, (int)row_type.size());
            row_type.push_back(t);
            cl.push_back(t == 'L' ? -INFINITY : 0.0);
            cu.push_back(t == 'G' ? INFINITY : 0.0);
        }
    }
    int column(const std::string &name) {
        auto it = cols.find(name);
        if (it != cols.end()) return it->second;
        const int j = (int)c.size();
        cols.emplace(name, j);
        c.push_back(0.0);
        lb.push_back(0.0);
        ub.push_back(INFINITY);
        return j;
    }
    bool is_obj(const std::string &r) const { return has_obj && r == obj_row; }
};

qp_problem_t *read_mps(const char *filename) {
    LineReader in(filename);
    if (!in.ok()) {
        std::fprintf(stderr, "ERROR: could not open '%s'.\n", filename);
        return nullptr;
    }
    Mps s;
    Mps::Sec sec = Mps::NONE;
    std::string line;
    bool error = false;
    while (!error && in.next(line)) {
        if (line.empty() || line[0] == '*' || line[0] == '\n' || line[0] == '\r') continue;
        std::vector<std::string> tok;
        {
            size_t i = 0;
            while (i < line.size() && tok.size() < 6) {
                while (i < line.size() && std::strchr(" \t\n\r", line[i])) ++i;
                if (i >= line.size()) break;
                size_t j = i;
                while (j < line.size() && !std::strchr(" \t\n\r", line[j])) ++j;
                tok.push_back(line.substr(i, j - i));
                i = j;
            }
        }
        const int nt = (int)tok.size();
        if (nt == 0) continue;
        if (nt == 1 && std::isalpha((unsigned char)tok[0][0])) {
            static const std::pair<const char *, Mps::Sec> names[] = {
                {"ROWS", Mps::ROWS},         {"COLUMNS", Mps::COLUMNS}, {"RHS", Mps::RHS},
                {"RANGES", Mps::RANGES},     {"BOUNDS", Mps::BOUNDS},   {"OBJSENSE", Mps::OBJSENSE},
                {"QUADOBJ", Mps::QUADOBJ},   {"QMATRIX", Mps::QMATRIX}, {"ENDATA", Mps::ENDATA}};
            Mps::Sec next = Mps::NONE;
            for (const auto &p : names)
                if (tok[0] == p.first) next = p.second;
            if (sec == Mps::ROWS && next != Mps::ROWS) s.finish_rows();
            sec = next;
            if (sec == Mps::ENDATA) break;
            continue;
        }
        if (nt == 2 && tok[0] == "QCMATRIX") {
            if (sec == Mps::ROWS) s.finish_rows();
            sec = Mps::QCMATRIX;
            s.qc_cur = -1;
            s.qc_obj = false;
            if (s.is_obj(tok[1])) {
                s.qc_obj = true;
            } else {
                const int r = s.row_of(tok[1]);
                if (r < 0) {
                    std::fprintf(stderr, "Warning: QCMATRIX references unknown row '%s'. Skipping.\n", tok[1].c_str());
                } else {
                    for (size_t i = 0; i < s.qcs.size(); ++i)
                        if (s.qcs[i].row == r) s.qc_cur = (int)i;
                    if (s.qc_cur < 0) {
                        s.qcs.push_back({r, Coo()});
                        s.qc_cur = (int)s.qcs.size() - 1;
                    }
                }
            }
            continue;
        }
        switch (sec) {
        case Mps::OBJSENSE:
            if (tok[0] == "MAX" || tok[0] == "MAXIMIZE") s.maximize = true;
            break;
        case Mps::ROWS:
            if (nt >= 2) s.raw_rows.push_back({tok[0][0], tok[1]});
            break;
        case Mps::COLUMNS: {
            if (nt < 2 || tok[1] == "'MARKER'") break;
            int first = 0;
            if (nt % 2 != 0) {
                s.cur_col = tok[0];
                s.have_col = true;
                first = 1;
            } else if (!s.have_col) {
                std::fprintf(stderr, "ERROR: Column data found before any column name was defined.\n");
                error = true;
                break;
            }
            const int j = s.column(s.cur_col);
            for (int i = first; i + 1 < nt; i += 2) {
                const double v = std::atof(tok[i + 1].c_str());
                if (s.is_obj(tok[i])) s.c[j] += v;
                else {
                    const int r = s.row_of(tok[i]);
                    if (r >= 0) s.A.add(r, j, v);
                }
            }
            break;
        }
        case Mps::RHS:
            for (int i = 1; i + 1 < nt; i += 2) {
                const double v = std::atof(tok[i + 1].c_str());
                if (s.is_obj(tok[i])) {
                    s.obj_const = -v;
                    continue;
                }
                const int r = s.row_of(tok[i]);
                if (r < 0) continue;
                if (s.row_type[r] == 'L') s.cu[r] = v;
                else if (s.row_type[r] == 'G') s.cl[r] = v;
                else s.cl[r] = s.cu[r] = v;
            }
            break;
        case Mps::RANGES:
            for (int i = 1; i + 1 < nt; i += 2) {
                const double R = std::atof(tok[i + 1].c_str());
                const int r = s.row_of(tok[i]);
                if (r < 0) continue;
                const char t = s.row_type[r];
                const double rhs = t == 'L' ? s.cu[r] : s.cl[r];
                if (t == 'G') s.cu[r] = rhs + std::fabs(R);
                else if (t == 'L') s.cl[r] = rhs - std::fabs(R);
                else if (R >= 0) s.cu[r] = rhs + R;
                else s.cl[r] = rhs + R;
            }
            break;
        case Mps::BOUNDS: {
            if (nt < 3) break;
            const int j = s.col_of(tok[2]);
            if (j < 0) break;
            const double v = nt > 3 ? std::atof(tok[3].c_str()) : 0.0;
            const std::string &t = tok[0];
            if (t == "LO") s.lb[j] = v;
            else if (t == "UP") s.ub[j] = v;
            else if (t == "FX") s.lb[j] = s.ub[j] = v;
            else if (t == "FR") {
                s.lb[j] = -INFINITY;
                s.ub[j] = INFINITY;
            } else if (t == "MI") s.lb[j] = -INFINITY;
            else if (t == "PL") s.ub[j] = INFINITY;
            else if (t == "BV") {
                s.lb[j] = 0.0;
                s.ub[j] = 1.0;
            }
            break;
        }
        case Mps::QUADOBJ:
        case Mps::QMATRIX: {
            if (nt < 3) break;
            const double v = std::atof(tok[2].c_str());
            if (v == 0.0) break;
            const int a = s.col_of(tok[0]), b = s.col_of(tok[1]);
            if (a < 0 || b < 0) {
                std::fprintf(stderr, "Warning: Variable '%s' or '%s' not found in COLUMNS. Skipping Q-entry.\n",
                             tok[0].c_str(), tok[1].c_str());
                break;
            }
            s.Q.add(a, b, v);
            if (a != b && sec == Mps::QUADOBJ) s.Q.add(b, a, v);  // QUADOBJ lists one triangle
            break;
        }
        case Mps::QCMATRIX: {
            if (nt < 3 || (s.qc_cur < 0 && !s.qc_obj)) break;
            const double v = std::atof(tok[2].c_str());
            if (v == 0.0) break;
            const int a = s.col_of(tok[0]), b = s.col_of(tok[1]);
            if (a < 0 || b < 0) {
                std::fprintf(stderr, "Warning: QCMATRIX references unknown variable '%s' or '%s'. Skipping entry.\n",
                             tok[0].c_str(), tok[1].c_str());
                break;
            }
            (s.qc_obj ? s.Q : s.qcs[s.qc_cur].m).add(a, b, v);
            break;
        }
        default:
            break;
        }
    }
    if (error) {
        std::fprintf(stderr, "ERROR: Failed to parse MPS file.\n");
        return nullptr;
    }
    s.finish_rows();
    const int n = (int)s.c.size(), m = (int)s.row_type.size();
    const double sgn = s.maximize ? -1.0 : 1.0;
    qp_problem_t *p = ccalloc<qp_problem_t>(1);
    try {
        p->num_variables = n;
        p->num_constraints = m;
        p->num_original_variables = n;
        p->affine_cone_offset = m > 0 ? ccalloc<double>(m) : nullptr;
        p->objective_constant = sgn * s.obj_const;
        for (double &v : s.c) v *= sgn;
        for (double &v : s.Q.v) v *= sgn;
        p->objective_vector = copy_out(s.c);
        p->variable_lower_bound = copy_out(s.lb);
        p->variable_upper_bound = copy_out(s.ub);
        p->constraint_lower_bound = copy_out(s.cl);
        p->constraint_upper_bound = copy_out(s.cu);
        p->constraint_matrix = coo_rows_stable(s.A, m);
        p->constraint_matrix_num_nonzeros = (int)s.A.nnz();
        p->objective_sparse_matrix = coo_rows_stable(s.Q, n);
        p->objective_sparse_matrix_num_nonzeros = (int)s.Q.nnz();
        p->objective_lowrank_matrix = ccalloc<CsrComponent>(1);
        p->objective_lowrank_matrix->row_ptr = ccalloc<int>(1);
        std::vector<const Mps::Qc *> kept;
        for (const Mps::Qc &q : s.qcs)
            if (q.m.nnz()) kept.push_back(&q);
        if (!kept.empty()) {
            const int K = (int)kept.size();
            p->num_quadratic_constraints = K;
            p->quadratic_constraint_row_indices = ccalloc<int>(K);
            p->quadratic_constraint_matrices = ccalloc<CsrComponent *>(K);
            p->quadratic_constraint_matrix_num_nonzeros = ccalloc<int>(K);
            for (int k = 0; k < K; ++k) {
                Coo q = kept[k]->m;
                for (double &v : q.v) v *= sgn;
                p->quadratic_constraint_row_indices[k] = kept[k]->row;
                p->quadratic_constraint_matrix_num_nonzeros[k] = (int)q.nnz();
                p->quadratic_constraint_matrices[k] = coo_rows_stable(q, n);
            }
        }
    } catch (const std::exception &e) {
        std::fprintf(stderr, "ERROR: %s\n", e.what());
        qp_problem_free(p);
        return nullptr;
    }
    return p;
}

// ---------------------------------------------------------------- CBF

enum CbfCone { kFree, kZero, kNonneg, kNonpos, kSoc, kRsoc, kExp };

bool cbf_cone(const std::string &t, CbfCone &out) {
    static const std::pair<const char *, CbfCone> names[] = {{"F", kFree}, {"L=", kZero}, {"L+", kNonneg},
                                                            {"L-", kNonpos}, {"Q", kSoc}, {"QR", kRsoc},
                                                            {"EXP", kExp}};
    for (const auto &p : names)
        if (t == p.first) {
            out = p.second;
            return true;
        }
    return false;
}
bool linear_cone(CbfCone c) { return c <= kNonpos; }

// Position inside the library's block of component `idx` of a CBF cone of
// dimension dim: CBF Q is (t, v...) -> [v..., w, z] with z = t and w the last
// CBF slot; QR is (s, t, v...) -> [v..., s, t]; EXP (x0, x1, x2) -> (x2, x1, x0).
int cbf_slot(CbfCone c, int dim, int idx) {
    if (c == kSoc) return idx == 0 ? dim - 1 : (idx == dim - 1 ? dim - 2 : idx - 1);
    if (c == kRsoc) return idx == 0 ? dim - 2 : (idx == 1 ? dim - 1 : idx - 2);
    return 2 - idx;
}
int cbf_vdim(CbfCone c, int dim) { return (c == kSoc || c == kRsoc) ? dim - 2 : 1; }
cone_type_t cbf_type(CbfCone c) {
    return c == kSoc ? CONE_STANDARD_SOC : (c == kRsoc ? CONE_ROTATED_SOC : CONE_EXPONENTIAL);
}

struct CbfBlock {
    CbfCone cone;
    int dim, start;
};

class CbfReader {
  public:
    explicit CbfReader(const char *path) : in_(path) {}
    bool ok() const { return in_.ok(); }
    // next non-blank, non-comment line, trailing blanks removed
    bool next(std::string &s) {
        while (in_.next(s)) {
            size_t i = 0;
            while (i < s.size() && (s[i] == ' ' || s[i] == '\t')) ++i;
            if (i >= s.size() || s[i] == '\n' || s[i] == '\r' || s[i] == '#') continue;
            while (!s.empty() && std::strchr("\n\r \t", s.back())) s.pop_back();
            return true;
        }
        return false;
    }

  private:
    LineReader in_;
};

bool cbf_blocks(CbfReader &r, int count, int &total, std::vector<CbfBlock> &out, const char *ctx) {
    total = 0;
    std::string ln;
    for (int i = 0; i < count; ++i) {
        char name[16];
        int dim = 0;
        if (!r.next(ln) || std::sscanf(ln.c_str(), "%15s %d", name, &dim) != 2 || dim <= 0) {
            std::fprintf(stderr, "[cbf] %s: bad cone line '%s'\n", ctx, ln.c_str());
            return false;
        }
        CbfCone c;
        if (!cbf_cone(name, c)) {
            std::fprintf(stderr, "[cbf] %s: unsupported cone '%s' (need L=/L+/L-/F/Q/QR/EXP)\n", ctx, name);
            return false;
        }
        if ((c == kSoc && dim < 2) || (c == kRsoc && dim < 3) || (c == kExp && dim != 3)) {
            std::fprintf(stderr, "[cbf] %s: bad dimension %d for cone '%s'\n", ctx, dim, name);
            return false;
        }
        out.push_back({c, dim, total});
        total += dim;
    }
    return true;
}

qp_problem_t *read_cbf(const char *filename) {
    CbfReader r(filename);
    if (!r.ok()) {
        std::fprintf(stderr, "[cbf] could not open '%s'\n", filename);
        return nullptr;
    }
    std::vector<CbfBlock> vb, cb;
    int n = 0, m = 0;
    bool have_var = false, have_con = false, negate = false;
    std::vector<double> c, b;
    double obj_const = 0.0;
    Coo A, Q;
    std::string ln;
    bool ok = true;
    auto count_line = [&](int &k) { return r.next(ln) && std::sscanf(ln.c_str(), "%d", &k) == 1 && k >= 0; };
    while (ok && r.next(ln)) {
        char kw[32];
        if (std::sscanf(ln.c_str(), "%31s", kw) != 1) {
            ok = false;
            break;
        }
        const std::string k = kw;
        int cnt = 0;
        if (k == "VER") {
            ok = r.next(ln) && std::sscanf(ln.c_str(), "%d", &cnt) == 1;
        } else if (k == "OBJSENSE") {
            char sense[16];
            ok = r.next(ln) && std::sscanf(ln.c_str(), "%15s", sense) == 1 &&
                 (std::strcmp(sense, "MIN") == 0 || std::strcmp(sense, "MAX") == 0);
            if (ok) negate = std::strcmp(sense, "MAX") == 0;
        } else if (k == "VAR" || k == "CON") {
            const bool var = k == "VAR";
            int dim = 0, nb = 0, total = 0;
            if ((var ? have_var : have_con) || !r.next(ln) || std::sscanf(ln.c_str(), "%d %d", &dim, &nb) != 2 ||
                dim < 0 || nb < 0 || !cbf_blocks(r, nb, total, var ? vb : cb, kw) || total != dim) {
                std::fprintf(stderr, "[cbf] bad %s section\n", kw);
                ok = false;
                break;
            }
            if (var) {
                have_var = true;
                n = dim;
                c.assign(n, 0.0);
            } else {
                have_con = true;
                m = dim;
                b.assign(m, 0.0);
            }
        } else if (k == "OBJACOORD" || k == "OBJQCOORD" || k == "ACOORD" || k == "BCOORD") {
            const bool need_con = k == "ACOORD" || k == "BCOORD";
            if (!count_line(cnt) || !have_var || (need_con && !have_con)) {
                ok = false;
                break;
            }
            for (int e = 0; e < cnt && ok; ++e) {
                int i = 0, j = 0;
                double v = 0.0;
                if (!r.next(ln)) ok = false;
                else if (k == "OBJACOORD") {
                    ok = std::sscanf(ln.c_str(), "%d %lf", &j, &v) == 2 && j >= 0 && j < n;
                    if (ok) c[j] += v;
                } else if (k == "BCOORD") {
                    ok = std::sscanf(ln.c_str(), "%d %lf", &i, &v) == 2 && i >= 0 && i < m;
                    if (ok) b[i] += v;
                } else if (k == "ACOORD") {
                    ok = std::sscanf(ln.c_str(), "%d %d %lf", &i, &j, &v) == 3 && i >= 0 && i < m && j >= 0 && j < n;
                    if (ok) A.add(i, j, v);
                } else {  // OBJQCOORD: objective 1/2 x'Qx, one triangle listed
                    ok = std::sscanf(ln.c_str(), "%d %d %lf", &i, &j, &v) == 3 && i >= 0 && i < n && j >= 0 && j < n;
                    if (ok) {
                        Q.add(i, j, v);
                        if (i != j) Q.add(j, i, v);
                    }
                }
            }
            if (!ok) std::fprintf(stderr, "[cbf] bad %s entry '%s'\n", kw, ln.c_str());
        } else if (k == "OBJBCOORD") {
            double v = 0.0;
            ok = r.next(ln) && std::sscanf(ln.c_str(), "%lf", &v) == 1;
            obj_const += v;
        } else if (k == "CHANGE") {
            while (r.next(ln)) {
            }  // modification sections: the base problem is complete
            break;
        } else {
            // INT (integers), PSDVAR, PSDCON, OBJFCOORD, FCOORD, HCOORD, DCOORD
            // and anything unknown are not supported
            std::fprintf(stderr, "[cbf] unsupported section '%s'\n", kw);
            ok = false;
        }
    }
    if (!ok || !have_var || vb.empty()) {
        if (ok) std::fprintf(stderr, "[cbf] no VAR section\n");
        return nullptr;
    }
    // columns: linear blocks first (file order), then cone blocks
    std::vector<int> col_of((size_t)n), lp_off(vb.size(), -1), cone_off(vb.size(), -1);
    int cols = 0;
    for (size_t i = 0; i < vb.size(); ++i)
        if (linear_cone(vb[i].cone)) {
            lp_off[i] = cols;
            cols += vb[i].dim;
        }
    for (size_t i = 0; i < vb.size(); ++i)
        if (!linear_cone(vb[i].cone)) {
            cone_off[i] = cols;
            cols += vb[i].dim;
        }
    for (size_t i = 0; i < vb.size(); ++i)
        for (int j = 0; j < vb[i].dim; ++j)
            col_of[vb[i].start + j] = linear_cone(vb[i].cone) ? lp_off[i] + j
                                                              : cone_off[i] + cbf_slot(vb[i].cone, vb[i].dim, j);
    // rows: constraint blocks in order, cone components permuted within a block
    std::vector<int> row_of((size_t)m), block_row((size_t)cb.size());
    int rc = 0;
    for (size_t i = 0; i < cb.size(); ++i) {
        block_row[i] = rc;
        for (int j = 0; j < cb[i].dim; ++j)
            row_of[cb[i].start + j] = linear_cone(cb[i].cone) ? rc + j : rc + cbf_slot(cb[i].cone, cb[i].dim, j);
        rc += cb[i].dim;
    }
    qp_problem_t *p = ccalloc<qp_problem_t>(1);
    try {
        const double sgn = negate ? -1.0 : 1.0;
        p->num_variables = cols;
        p->num_constraints = m;
        p->num_original_variables = cols;
        p->objective_constant = obj_const;
        p->affine_cone_offset = m > 0 ? ccalloc<double>(m) : nullptr;
        p->objective_vector = ccalloc<double>(cols);
        for (int i = 0; i < n; ++i) p->objective_vector[col_of[i]] = sgn * c[i];
        p->variable_lower_bound = cmalloc<double>(cols);
        p->variable_upper_bound = cmalloc<double>(cols);
        for (int j = 0; j < cols; ++j) {
            p->variable_lower_bound[j] = -INFINITY;
            p->variable_upper_bound[j] = INFINITY;
        }
        std::vector<int> vstart, vvd;
        std::vector<cone_type_t> vtype;
        for (size_t i = 0; i < vb.size(); ++i) {
            const int base = lp_off[i];
            for (int j = 0; base >= 0 && j < vb[i].dim; ++j) {
                if (vb[i].cone == kNonneg || vb[i].cone == kZero) p->variable_lower_bound[base + j] = 0.0;
                if (vb[i].cone == kNonpos || vb[i].cone == kZero) p->variable_upper_bound[base + j] = 0.0;
            }
            if (!linear_cone(vb[i].cone)) {
                vstart.push_back(cone_off[i]);
                vvd.push_back(cbf_vdim(vb[i].cone, vb[i].dim));
                vtype.push_back(cbf_type(vb[i].cone));
            }
        }
        p->cones.num_cones = (int)vstart.size();
        if (!vstart.empty()) {
            p->cones.start_idx = copy_out(vstart);
            p->cones.v_dim = copy_out(vvd);
            p->cones.type = copy_out(vtype);
        }
        std::vector<int> astart, avd;
        std::vector<cone_type_t> atype;
        for (size_t i = 0; i < cb.size(); ++i)
            if (!linear_cone(cb[i].cone)) {
                astart.push_back(block_row[i]);
                avd.push_back(cbf_vdim(cb[i].cone, cb[i].dim));
                atype.push_back(cbf_type(cb[i].cone));
            }
        p->affine_cones.num_cones = (int)astart.size();
        if (!astart.empty()) {
            p->affine_cones.start_idx = copy_out(astart);
            p->affine_cones.v_dim = copy_out(avd);
            p->affine_cones.type = copy_out(atype);
        }
        Coo Am;
        for (size_t e = 0; e < A.nnz(); ++e) Am.add(row_of[A.r[e]], col_of[A.c[e]], A.v[e]);
        int nnz = 0;
        p->constraint_matrix = coo_rows_merged(Am, m, &nnz);
        p->constraint_matrix_num_nonzeros = nnz;
        // A x + b in K: linear rows get bounds -b; cone rows keep b as the offset
        p->constraint_lower_bound = cmalloc<double>(std::max(m, 1));
        p->constraint_upper_bound = cmalloc<double>(std::max(m, 1));
        for (int i = 0; i < m; ++i) {
            p->constraint_lower_bound[i] = -INFINITY;
            p->constraint_upper_bound[i] = INFINITY;
        }
        for (const CbfBlock &blk : cb)
            for (int j = 0; j < blk.dim; ++j) {
                const int src = blk.start + j, row = row_of[src];
                if (!linear_cone(blk.cone)) {
                    p->affine_cone_offset[row] = b[src];
                    continue;
                }
                const double z = -b[src];
                if (blk.cone == kZero || blk.cone == kNonneg) p->constraint_lower_bound[row] = z;
                if (blk.cone == kZero || blk.cone == kNonpos) p->constraint_upper_bound[row] = z;
            }
        // a singleton equality row fixing the radius slot of a variable SOC
        // block (one consistent value >= 0) marks that slot fixed
        if (p->cones.num_cones > 0 && m > 0) {
            std::vector<char> radius((size_t)cols, 0), state((size_t)cols, 0);
            std::vector<double> value((size_t)cols, 0.0);
            for (int k = 0; k < p->cones.num_cones; ++k)
                if (p->cones.type[k] == CONE_STANDARD_SOC) {
                    const int z = p->cones.start_idx[k] + p->cones.v_dim[k] + 1;
                    if (z >= 0 && z < cols) radius[z] = 1;
                }
            const CsrComponent *M = p->constraint_matrix;
            for (int i = 0; i < m; ++i) {
                const double lo = p->constraint_lower_bound[i], hi = p->constraint_upper_bound[i];
                if (M->row_ptr[i + 1] - M->row_ptr[i] != 1 || !std::isfinite(lo) || lo != hi) continue;
                const int j = M->col_ind[M->row_ptr[i]];
                const double a = M->val[M->row_ptr[i]];
                if (!radius[j] || !std::isfinite(a) || a == 0.0) continue;
                const double v = (lo - p->affine_cone_offset[i]) / a;
                if (!std::isfinite(v) || v < 0.0) continue;
                if (state[j] == 0) {
                    state[j] = 1;
                    value[j] = v;
                } else if (state[j] == 1 && std::fabs(value[j] - v) > 1e-12 * (1.0 + std::fmax(std::fabs(value[j]),
                                                                                               std::fabs(v)))) {
                    state[j] = 2;
                }
            }
            if (std::count(state.begin(), state.end(), 1) > 0) {
                p->cones.fixed_mask_size = cols;
                p->cones.is_fixed = ccalloc<char>(cols);
                p->primal_start = ccalloc<double>(cols);
                for (int j = 0; j < cols; ++j)
                    if (state[j] == 1) {
                        p->cones.is_fixed[j] = 1;
                        p->primal_start[j] = value[j];
                    }
            }
        }
        if (Q.nnz()) {
            Coo Qm;
            for (size_t e = 0; e < Q.nnz(); ++e) Qm.add(col_of[Q.r[e]], col_of[Q.c[e]], sgn * Q.v[e]);
            int qn = 0;
            p->objective_sparse_matrix = coo_rows_merged(Qm, cols, &qn);
            p->objective_sparse_matrix_num_nonzeros = qn;
        }
    } catch (const std::exception &e) {
        std::fprintf(stderr, "[cbf] %s\n", e.what());
        qp_problem_free(p);
        return nullptr;
    }
    return p;
}

}  // namespace
}  // namespace pdhcg

extern "C" {

// reference include/mps_parser.h (src/mps_parser.c:420)
qp_problem_t *read_mps_file(const char *filename) {
    if (!filename) return nullptr;
    try {
        return pdhcg::read_mps(filename);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "ERROR: %s\n", e.what());
        return nullptr;
    }
}

// reference include/cbf_parser.h (src/cbf_parser.c:1178)
qp_problem_t *read_cbf_file(const char *filename) {
    if (!filename) return nullptr;
    try {
        return pdhcg::read_cbf(filename);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "[cbf] %s\n", e.what());
        return nullptr;
    }
}

}  // extern "C"
