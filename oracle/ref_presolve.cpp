// C entry points over the reference's C++ presolve -- TEST INFRASTRUCTURE ONLY.
//
// Built by `make -C oracle ref` together with the reference's own sources
// (src/lp/Presolve.cpp, src/lp/LpProblem.cpp, src/sparse/*.cpp,
// src/io/MpsReader.cpp, compiled where they lie) into
// oracle/_ref/libsihps_presolve_ref.so.  It only converts between plain
// arrays and the reference's LpProblem / PresolveResult so that the tests
// can run the reference's presolve as the checker of sk_gpu_presolve.
#include "io/MpsReader.hpp"
#include "lp/LpProblem.hpp"
#include "lp/Presolve.hpp"
#include "sparse/Triplet.hpp"

#include "sankhya_gpu.h"

#include <cstdlib>
#include <cstring>
#include <exception>
#include <vector>

namespace {

template <typename T>
T *copy_out(const std::vector<T> &v)
{
    T *p = static_cast<T *>(std::malloc(sizeof(T) * (v.size() + 1)));
    if (p && !v.empty()) std::memcpy(p, v.data(), sizeof(T) * v.size());
    return p;
}

void export_problem(const sihps::LpProblem &p, sk_gpu_lp_problem *out)
{
    std::memset(out, 0, sizeof(*out));
    out->rows = p.n_rows();
    out->cols = p.n_cols();
    out->nnz = p.A.nnz();
    out->row_ptr = copy_out(std::vector<int>(p.A.row_ptr(), p.A.row_ptr() + p.n_rows() + 1));
    out->col_idx = copy_out(std::vector<int>(p.A.col_idx(), p.A.col_idx() + p.A.nnz()));
    out->values = copy_out(std::vector<double>(p.A.values(), p.A.values() + p.A.nnz()));
    out->obj = copy_out(p.obj);
    out->rhs = copy_out(p.rhs);
    out->row_types = copy_out(p.row_types);
    out->lower = copy_out(p.lower);
    out->upper = copy_out(p.upper);
    out->slack_lower = copy_out(p.slack_lower);
    out->slack_upper = copy_out(p.slack_upper);
}

sihps::LpProblem import_problem(const sk_gpu_lp_problem *in)
{
    sihps::LpProblem p;
    const int m = in->rows, n = in->cols, nnz = in->nnz;
    p.A = sihps::CSRMatrix(m, n, std::vector<std::int32_t>(in->row_ptr, in->row_ptr + m + 1),
                           std::vector<std::int32_t>(in->col_idx, in->col_idx + nnz),
                           std::vector<double>(in->values, in->values + nnz));
    p.obj.assign(in->obj, in->obj + n);
    p.lower.assign(in->lower, in->lower + n);
    p.upper.assign(in->upper, in->upper + n);
    p.rhs.assign(in->rhs, in->rhs + m);
    p.slack_lower.assign(in->slack_lower, in->slack_lower + m);
    p.slack_upper.assign(in->slack_upper, in->slack_upper + m);
    if (in->row_types) p.row_types.assign(in->row_types, in->row_types + m);
    else p.row_types.assign(static_cast<std::size_t>(m), 'L');
    return p;
}

} // namespace

extern "C" {

// read_mps_file + lp_problem_from_mps (src/io/MpsReader.hpp, src/lp/LpProblem.cpp:30)
int refps_read_mps(const char *path, sk_gpu_lp_problem *out)
{
    try {
        const sihps::MpsModel model = sihps::read_mps_file(path);
        export_problem(sihps::lp_problem_from_mps(model), out);
        return 0;
    } catch (const std::exception &) {
        return -1;
    }
}

// presolve (src/lp/Presolve.hpp:92) into the layout of sk_gpu_presolve_result
int refps_presolve(const sk_gpu_lp_problem *in, int max_passes, sk_gpu_presolve_result *out)
{
    try {
        std::memset(out, 0, sizeof(*out));
        const sihps::LpProblem p = import_problem(in);
        const sihps::PresolveResult r = sihps::presolve(p, max_passes);
        out->status = static_cast<int>(r.status);
        out->reason = r.reason;
        out->reason_row = r.reason_row;
        out->reason_col = r.reason_col;
        out->reason_pass = r.reason_pass;
        out->original_n_rows = r.original_n_rows;
        out->original_n_cols = r.original_n_cols;
        out->removed_rows = r.removed_rows();
        out->removed_cols = r.removed_cols();
        out->kept_columns = copy_out(r.kept_columns);
        out->kept_rows = copy_out(r.kept_rows);
        out->fixed_value = copy_out(r.fixed_value);
        out->column_removed = copy_out(r.column_removed);
        if (r.status == sihps::PresolveStatus::OK) export_problem(r.reduced, &out->reduced);
        return 0;
    } catch (const std::exception &) {
        return -1;
    }
}

// postsolve (src/lp/Presolve.hpp:95)
int refps_postsolve(const sk_gpu_lp_problem *in, int max_passes, const double *reduced_x, int reduced_len,
                    double *x)
{
    try {
        const sihps::LpProblem p = import_problem(in);
        const sihps::PresolveResult r = sihps::presolve(p, max_passes);
        const std::vector<double> full =
            sihps::postsolve(r, std::vector<double>(reduced_x, reduced_x + reduced_len));
        std::memcpy(x, full.data(), sizeof(double) * full.size());
        return 0;
    } catch (const std::exception &) {
        return -1;
    }
}

void refps_problem_free(sk_gpu_lp_problem *p)
{
    std::free(p->row_ptr); std::free(p->col_idx); std::free(p->values); std::free(p->obj);
    std::free(p->rhs); std::free(p->row_types); std::free(p->lower); std::free(p->upper);
    std::free(p->slack_lower); std::free(p->slack_upper);
    std::memset(p, 0, sizeof(*p));
}

void refps_result_free(sk_gpu_presolve_result *r)
{
    refps_problem_free(&r->reduced);
    std::free(r->kept_columns); std::free(r->kept_rows); std::free(r->fixed_value);
    std::free(r->column_removed);
    std::memset(r, 0, sizeof(*r));
}

} // extern "C"
