// cbf.cpp — reader of the Conic Benchmark Format (.cbf) into the canonical
// layout of include/hprsocp.h.  It follows read_cbf / build_from_cbf of the
// reference (src/utils/cbf_io.jl:493-750, 899-1131; src/utils/builders.jl:24-130):
//   * constraint rows are reordered L= | L+ and L- | Q; L- rows are negated;
//     F (free) constraint rows are dropped;
//   * the file's form A x + b in K becomes A x - rhs in K with rhs = -b
//     (L- rows: rhs = +b after the negation);
//   * variables are reordered linear (F, L+, L-, L=) | Q, with bounds
//     F: (-inf, inf), L+: [0, inf), L-: (-inf, 0], L=: [0, 0];
//   * equality rows get AL = AU = rhs, inequality rows AL = rhs, AU = +inf,
//     SOC rows carry soc_rhs;
//   * OBJACOORD entries assign c, OBJBCOORD the constant, SENSE/OBJSENSE max
//     negates both; A entries of value 0 are dropped and duplicates summed;
//   * PSD, matrix-variable and integer sections must be empty.
#include <errno.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <utility>
#include <vector>

#include "../../include/hprsocp.h"

struct hprsocp_cbf {
    int32_t m = 0, n = 0, neq = 0, nineq = 0, n1 = 0;
    std::vector<int32_t> rp, ci, con_idx, var_idx;
    std::vector<double> v, c, AL, AU, l, u, soc_rhs;
    double obj_const = 0.0;
};

namespace {

std::string g_cbf_error;

struct CbfErr {
    std::string msg;
};

// Line cursor over the file text: next() returns the next line trimmed of
// spaces, tabs and CR; data_line() also skips empty lines and '#' comments.
struct Lines {
    const char *p, *end;
    bool next(std::string &out) {
        if (p >= end) return false;
        const char *e = (const char *)memchr(p, '\n', (size_t)(end - p));
        if (!e) e = end;
        const char *a = p, *b = e;
        while (a < b && (*a == ' ' || *a == '\t' || *a == '\r')) ++a;
        while (b > a && (b[-1] == ' ' || b[-1] == '\t' || b[-1] == '\r')) --b;
        out.assign(a, (size_t)(b - a));
        p = e < end ? e + 1 : end;
        return true;
    }
    bool data_line(std::string &out) {
        while (next(out))
            if (!out.empty() && out[0] != '#') return true;
        return false;
    }
    std::string need() {
        std::string s;
        if (!next(s)) throw CbfErr{"Invalid CBF file: unexpected end of file."};
        return s;
    }
};

long long parse_int(const char *&s) {
    char *e = nullptr;
    errno = 0;
    const long long v = strtoll(s, &e, 10);
    if (e == s || errno) throw CbfErr{"Invalid CBF line: expected integer token."};
    s = e;
    return v;
}
double parse_double(const char *&s) {
    char *e = nullptr;
    const double v = strtod(s, &e);
    if (e == s) throw CbfErr{"Invalid CBF line: expected floating-point token."};
    s = e;
    return v;
}
std::pair<std::string, long long> parse_group(const std::string &line) {
    const char *s = line.c_str();
    while (*s == ' ' || *s == '\t') ++s;
    const char *t = s;
    while (*t && *t != ' ' && *t != '\t') ++t;
    if (t == s) throw CbfErr{"Invalid CBF line: expected a cone tag and length."};
    std::string tag(s, (size_t)(t - s));
    const long long len = parse_int(t);
    return {tag, len};
}

bool supported(const std::string &t) { return t == "L+" || t == "L-" || t == "L=" || t == "Q" || t == "F"; }

hprsocp_cbf *read_impl(const char *path) {
    const std::string name(path ? path : "");
    const bool ext = name.size() >= 4 &&
                     (name.compare(name.size() - 4, 4, ".cbf") == 0 || name.compare(name.size() - 4, 4, ".CBF") == 0);
    if (!ext) throw CbfErr{"Unsupported file format. Please provide a .cbf file."};
    FILE *f = fopen(path, "rb");
    if (!f) throw CbfErr{"cannot open " + name};
    std::string text;
    char buf[1 << 16];
    size_t k;
    while ((k = fread(buf, 1, sizeof buf, f)) > 0) text.append(buf, k);
    fclose(f);

    std::vector<std::pair<std::string, long long>> con, var;
    long long nconstr = -1, nvar = -1;
    bool maximize = false;
    double objb = 0.0;
    std::vector<long long> ar, ac, bi, ci_;
    std::vector<double> av, bv, cv;
    Lines L{text.data(), text.data() + text.size()};
    std::string line;
    auto count = [&]() {
        const std::string s = L.need();
        const char *q = s.c_str();
        const long long c = parse_int(q);
        if (c < 0) throw CbfErr{"Invalid CBF file: negative entry count."};
        return c;
    };
    while (L.data_line(line)) {
        if (line == "CON" || line == "VAR") {
            const std::string s = L.need();
            const char *q = s.c_str();
            const long long tot = parse_int(q), ng = parse_int(q);
            auto &grp = line == "CON" ? con : var;
            grp.clear();
            for (long long j = 0; j < ng; ++j) grp.push_back(parse_group(L.need()));
            (line == "CON" ? nconstr : nvar) = tot;
        } else if (line == "ACOORD") {
            const long long nz = count();
            for (long long j = 0; j < nz; ++j) {
                const std::string s = L.need();
                const char *q = s.c_str();
                ar.push_back(parse_int(q));
                ac.push_back(parse_int(q));
                av.push_back(parse_double(q));
            }
        } else if (line == "BCOORD" || line == "OBJACOORD") {
            const long long nz = count();
            for (long long j = 0; j < nz; ++j) {
                const std::string s = L.need();
                const char *q = s.c_str();
                (line == "BCOORD" ? bi : ci_).push_back(parse_int(q));
                (line == "BCOORD" ? bv : cv).push_back(parse_double(q));
            }
        } else if (line == "OBJBCOORD") {
            const std::string s = L.need();
            const char *q = s.c_str();
            objb = parse_double(q);
        } else if (line == "SENSE" || line == "OBJSENSE") {
            std::string s = L.need();
            for (char &ch : s) ch = (char)tolower((unsigned char)ch);
            if (s == "max" || s == "maximize" || s == "maximise") maximize = true;
            else if (s == "min" || s == "minimize" || s == "minimise") maximize = false;
            else throw CbfErr{"Unsupported CBF objective sense: " + s + "."};
        } else if (line == "DCOORD" || line == "HCOORD" || line == "hCOORD" || line == "PSDCON" ||
                   line == "PSDVAR" || line == "INT" || line == "FCOORD" || line == "OBJFCOORD") {
            const long long nz = count();
            if (nz != 0) {
                if (line == "DCOORD" || line == "HCOORD" || line == "hCOORD" || line == "PSDCON")
                    throw CbfErr{"PSD constraints are not supported."};
                if (line == "FCOORD" || line == "OBJFCOORD" || line == "PSDVAR")
                    throw CbfErr{"Matrix variables are not supported."};
                throw CbfErr{"Integer variables are not supported."};
            }
        }
    }
    // layout (_cbf_prepare_layout, cbf_io.jl:493-749)
    if (con.empty()) throw CbfErr{"Invalid CBF file: missing CON section."};
    if (var.empty()) throw CbfErr{"Invalid CBF file: missing VAR section."};
    for (auto &g : con)
        if (!supported(g.first)) throw CbfErr{"Unsupported constraint cone type " + g.first + "."};
    for (auto &g : var)
        if (!supported(g.first)) throw CbfErr{"Unsupported variable cone type " + g.first + "."};
    long long m = 0, n = 0, neq = 0, nin = 0, nsoc = 0, nsv = 0;
    std::vector<long long> soc_len;
    for (auto &g : con) {
        if (g.second < 0) throw CbfErr{"Invalid CBF file: negative group length."};
        m += g.second;
        if (g.first == "L=") neq += g.second;
        else if (g.first == "L+" || g.first == "L-") nin += g.second;
        else if (g.first == "Q") nsoc += g.second, soc_len.push_back(g.second);
    }
    for (auto &g : var) {
        if (g.second < 0) throw CbfErr{"Invalid CBF file: negative group length."};
        n += g.second;
        if (g.first == "Q") nsv += g.second;
    }
    if (nconstr != -1 && nconstr != m) throw CbfErr{"CBF CON size mismatch."};
    if (nvar != -1 && nvar != n) throw CbfErr{"CBF VAR size mismatch."};
    if (m > INT32_MAX || n > INT32_MAX) throw CbfErr{"CBF file too large for 32-bit indices."};
    const long long total = neq + nin + nsoc;
    std::vector<int32_t> row_new((size_t)m, -1), col_new((size_t)n, -1);
    std::vector<int8_t> asign((size_t)m, 1), bsign((size_t)m, -1);
    {
        long long pe = 0, pi = neq, ps = neq + nin, r0 = 0;
        for (auto &g : con) {
            for (long long r = r0; r < r0 + g.second; ++r) {
                if (g.first == "L=") row_new[r] = (int32_t)pe++;
                else if (g.first == "L+") row_new[r] = (int32_t)pi++;
                else if (g.first == "L-") row_new[r] = (int32_t)pi++, asign[r] = -1, bsign[r] = 1;
                else if (g.first == "Q") row_new[r] = (int32_t)ps++;
            }
            r0 += g.second;
        }
    }
    hprsocp_cbf *M = new hprsocp_cbf();
    const long long n1 = n - nsv;
    M->m = (int32_t)total;
    M->n = (int32_t)n;
    M->neq = (int32_t)neq;
    M->nineq = (int32_t)nin;
    M->n1 = (int32_t)n1;
    M->l.assign((size_t)n, 0.0);
    M->u.assign((size_t)n, 0.0);
    M->var_idx.push_back((int32_t)n1);
    {
        long long pl = 0, pq = n1, c0 = 0;
        for (auto &g : var) {
            for (long long j = c0; j < c0 + g.second; ++j) {
                if (g.first == "Q") {
                    col_new[j] = (int32_t)pq;
                    M->l[pq] = -INFINITY;
                    M->u[pq] = INFINITY;
                    ++pq;
                } else {
                    col_new[j] = (int32_t)pl;
                    M->l[pl] = (g.first == "L+" || g.first == "L=") ? 0.0 : -INFINITY;
                    M->u[pl] = (g.first == "L-" || g.first == "L=") ? 0.0 : INFINITY;
                    ++pl;
                }
            }
            if (g.first == "Q") M->var_idx.push_back((int32_t)pq);
            c0 += g.second;
        }
    }
    M->con_idx.push_back((int32_t)(neq + nin));
    for (long long len : soc_len) M->con_idx.push_back(M->con_idx.back() + (int32_t)len);
    // objective
    M->c.assign((size_t)n, 0.0);
    for (size_t j = 0; j < ci_.size(); ++j) {
        if (ci_[j] < 0 || ci_[j] >= n) throw CbfErr{"OBJACOORD index out of range."};
        M->c[col_new[ci_[j]]] = cv[j];
    }
    M->obj_const = objb;
    if (maximize) {
        for (double &x : M->c) x = -x;
        M->obj_const = -M->obj_const;
    }
    // rhs
    std::vector<double> rhs((size_t)total, 0.0);
    for (size_t j = 0; j < bi.size(); ++j) {
        if (bi[j] < 0 || bi[j] >= m) throw CbfErr{"BCOORD index out of range."};
        const int32_t r = row_new[bi[j]];
        if (r >= 0) rhs[r] = (double)bsign[bi[j]] * bv[j];
    }
    M->AL.assign((size_t)total, -INFINITY);
    M->AU.assign((size_t)total, INFINITY);
    for (long long r = 0; r < neq; ++r) M->AL[r] = M->AU[r] = rhs[r];
    for (long long r = neq; r < neq + nin; ++r) M->AL[r] = rhs[r];
    M->soc_rhs.assign(rhs.begin() + (neq + nin), rhs.end());
    // A: mapped triplets -> CSR with rows sorted by column, duplicates summed, zeros dropped
    std::vector<std::pair<int32_t, std::pair<int32_t, double>>> T;
    T.reserve(av.size());
    for (size_t j = 0; j < av.size(); ++j) {
        if (ar[j] < 0 || ar[j] >= m || ac[j] < 0 || ac[j] >= n) throw CbfErr{"ACOORD index out of range."};
        const int32_t r = row_new[ar[j]];
        if (r < 0) continue;
        const double val = (double)asign[ar[j]] * av[j];
        if (val == 0.0) continue;
        T.push_back({r, {col_new[ac[j]], val}});
    }
    std::stable_sort(T.begin(), T.end(), [](const auto &a, const auto &b) {
        return a.first != b.first ? a.first < b.first : a.second.first < b.second.first;
    });
    M->rp.assign((size_t)total + 1, 0);
    for (size_t j = 0; j < T.size(); ++j) {
        const bool dup = !M->ci.empty() && j > 0 && T[j - 1].first == T[j].first &&
                         T[j - 1].second.first == T[j].second.first;
        if (dup) {
            M->v.back() += T[j].second.second;
        } else {
            M->ci.push_back(T[j].second.first);
            M->v.push_back(T[j].second.second);
            M->rp[T[j].first + 1]++;
        }
    }
    // drop entries whose duplicates cancelled (dropzeros!)
    {
        std::vector<int32_t> rp2((size_t)total + 1, 0), ci2;
        std::vector<double> v2;
        std::vector<int32_t> cnt(M->rp);
        for (long long r = 0; r < total; ++r) cnt[r + 1] += cnt[r];
        for (long long r = 0; r < total; ++r) {
            for (int32_t q = cnt[r]; q < cnt[r + 1]; ++q)
                if (M->v[q] != 0.0) {
                    ci2.push_back(M->ci[q]);
                    v2.push_back(M->v[q]);
                }
            rp2[r + 1] = (int32_t)ci2.size();
        }
        M->rp.swap(rp2);
        M->ci.swap(ci2);
        M->v.swap(v2);
    }
    return M;
}

}  // namespace

extern "C" {

int hprsocp_cbf_read(const char *path, hprsocp_cbf **out) {
    *out = nullptr;
    g_cbf_error.clear();
    try {
        *out = read_impl(path);
    } catch (const CbfErr &e) {
        g_cbf_error = e.msg;
        return HPRSOCP_ERR_INVALID;
    } catch (const std::bad_alloc &) {
        g_cbf_error = "out of memory";
        return HPRSOCP_ERR_ALLOC;
    }
    return 0;
}

const char *hprsocp_cbf_last_error(void) { return g_cbf_error.c_str(); }

int hprsocp_cbf_problem(const hprsocp_cbf *M, hprsocp_problem *p, int32_t *number_eq) {
    memset(p, 0, sizeof(*p));
    p->m = M->m;
    p->n = M->n;
    p->A_nnz = (int64_t)M->ci.size();
    p->A_rowptr = M->rp.data();
    p->A_colidx = M->ci.data();
    p->A_val = M->v.data();
    p->Q_kind = HPRSOCP_Q_CSR;
    p->Q_nnz = 0;
    p->c = M->c.data();
    p->AL = M->AL.data();
    p->AU = M->AU.data();
    p->l = M->l.data();
    p->u = M->u.data();
    p->soc_rhs = M->soc_rhs.data();
    p->number_linear_con = M->neq + M->nineq;
    p->number_lu_x = M->n1;
    p->n_soc_con = (int32_t)M->con_idx.size() - 1;
    p->soc_con_idx = M->con_idx.data();
    p->n_soc_var = (int32_t)M->var_idx.size() - 1;
    p->soc_var_idx = M->var_idx.data();
    p->obj_constant = M->obj_const;
    if (number_eq) *number_eq = M->neq;
    return 0;
}

void hprsocp_cbf_destroy(hprsocp_cbf *M) { delete M; }

}  // extern "C"
