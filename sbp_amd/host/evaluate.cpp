// Partition quality against a true assignment, on the host: the contingency table with a maximum-weight
// matching of communities (evaluate::hungarian, src/evaluate.cpp:134-210), pair-counting scores
// (calculate_f1_score, :9-59) and normalised mutual information (calculate_nmi, :61-96).
//
// The reference matches communities with the RoboJackets "hungarian" library (extern/hungarian/hungarian.cpp,
// vendored in its tree). The matching here is this project's own Hungarian method with shortest augmenting paths (Kuhn 1955; Munkres 1957; Jonker and Volgenant 1987) on the table padded to a square. Only
// the accuracy (the matched diagonal) depends on the matching; every maximum-weight matching gives the same
// value. F1, Rand index, precision, recall and NMI do not depend on the order of rows or columns.
#include <math.h>
#include <stdint.h>

#include <algorithm>
#include <limits>
#include <unordered_map>
#include <vector>

#include "../../include/sbp_amd.h"

namespace {

// Maximum-weight assignment of rows to columns of an n x n matrix (row-major).  Returns col_of[row].
std::vector<int32_t> max_weight_assignment(const std::vector<int64_t> &w, int32_t n) {
    int64_t top = 0;
    for (int64_t x : w) top = std::max(top, x);
    // minimise cost = top - w with potentials u (rows), v (columns); 1-based with column 0 as the root
    const int64_t INF = std::numeric_limits<int64_t>::max() / 4;
    std::vector<int64_t> u(n + 1, 0), v(n + 1, 0), minv(n + 1);
    std::vector<int32_t> p(n + 1, 0), way(n + 1, 0);
    std::vector<char> used(n + 1);
    for (int32_t i = 1; i <= n; ++i) {
        p[0] = i;
        int32_t j0 = 0;
        std::fill(minv.begin(), minv.end(), INF);
        std::fill(used.begin(), used.end(), 0);
        do {
            used[j0] = 1;
            const int32_t i0 = p[j0];
            int64_t delta = INF;
            int32_t j1 = 0;
            for (int32_t j = 1; j <= n; ++j) {
                if (used[j]) continue;
                const int64_t cur = (top - w[(size_t)(i0 - 1) * n + (j - 1)]) - u[i0] - v[j];
                if (cur < minv[j]) minv[j] = cur, way[j] = j0;
                if (minv[j] < delta) delta = minv[j], j1 = j;
            }
            for (int32_t j = 0; j <= n; ++j) {
                if (used[j]) u[p[j]] += delta, v[j] -= delta;
                else minv[j] -= delta;
            }
            j0 = j1;
        } while (p[j0] != 0);
        do {
            const int32_t j1 = way[j0];
            p[j0] = p[j1];
            j0 = j1;
        } while (j0);
    }
    std::vector<int32_t> col_of(n, 0);
    for (int32_t j = 1; j <= n; ++j)
        if (p[j]) col_of[p[j] - 1] = j - 1;
    return col_of;
}

}  // namespace

extern "C" {

int sbp_evaluate_table(int64_t num_vertices, int32_t nrows, int32_t ncols, const int32_t *table, sbp_evaluation *out) {
    if (num_vertices < 2 || nrows < 1 || ncols < 1 || !table || !out) return SBP_E_ARG;
    // calculate_f1_score (src/evaluate.cpp:9-59): pair counts over the table
    const double nv = (double)num_vertices;
    const double num_pairs = (nv * (nv - 1.0)) / 2.0;
    std::vector<double> rowsums(nrows, 0.0), colsums(ncols, 0.0);
    double cell_pairs = 0, cells_squared = 0, diagonal = 0;
    for (int32_t r = 0; r < nrows; ++r)
        for (int32_t c = 0; c < ncols; ++c) {
            const int64_t x = table[(size_t)r * ncols + c];
            rowsums[r] += (double)x;
            colsums[c] += (double)x;
            cell_pairs += (double)(x * (x - 1));
            cells_squared += (double)(x * x);
            if (r == c) diagonal += (double)x;
        }
    cell_pairs /= 2.0;
    double row_pairs = 0, col_pairs = 0, rows_squared = 0, cols_squared = 0;
    for (double s : rowsums) row_pairs += s * (s - 1), rows_squared += s * s;
    for (double s : colsums) col_pairs += s * (s - 1), cols_squared += s * s;
    row_pairs /= 2.0;
    col_pairs /= 2.0;
    double agree_apart = (double)(num_vertices * num_vertices) + cells_squared;
    agree_apart -= rows_squared + cols_squared;
    agree_apart /= 2.0;
    out->rand_index = (cell_pairs + agree_apart) / num_pairs;
    out->recall = cell_pairs / row_pairs;     // the reference's names: rows are the algorithm's communities
    out->precision = cell_pairs / col_pairs;
    out->f1_score = (2.0 * out->precision * out->recall) / (out->precision + out->recall);
    out->accuracy = diagonal / nv;
    // calculate_nmi (src/evaluate.cpp:61-96): the entropies use nat_log (logf of the probability rounded to
    // float, include/utils.hpp:228-238), the mutual information double log
    std::vector<double> pr(nrows, 0.0), pc(ncols, 0.0);
    for (int32_t r = 0; r < nrows; ++r)
        for (int32_t c = 0; c < ncols; ++c) {
            const double x = table[(size_t)r * ncols + c] / nv;
            pr[r] += x;
            pc[c] += x;
        }
    double hr = 0, hc = 0;
    for (double x : pr) hr += x * (double)logf((float)x);
    for (double x : pc) hc += x * (double)logf((float)x);
    hr = -hr;
    hc = -hc;
    double mi = 0;
    for (int32_t r = 0; r < nrows; ++r)
        for (int32_t c = 0; c < ncols; ++c) {
            const double joint = table[(size_t)r * ncols + c] / nv;
            if (joint == 0.0) continue;
            mi += joint * log(joint / (pr[r] * pc[c]));
        }
    out->nmi = mi / sqrt(hr * hc);
    out->found_blocks = nrows;
    out->true_blocks = ncols;
    return SBP_OK;
}

int sbp_evaluate(int32_t N, const int32_t *true_assignment, const int32_t *assignment, sbp_evaluation *out,
                 int32_t *table_out, int32_t *table_rows, int32_t *table_cols) {
    if (N < 2 || !true_assignment || !assignment || !out) return SBP_E_ARG;
    // true communities numbered in order of first appearance; -1 (unknown) leaves a vertex out of the table
    std::unordered_map<int32_t, int32_t> truth_id;
    std::vector<int32_t> truth(N, -1);
    int32_t num_true = 0, num_found = 0;
    for (int32_t i = 0; i < N; ++i) {
        if (assignment[i] < 0) return SBP_E_ARG;
        num_found = std::max(num_found, assignment[i] + 1);
        if (true_assignment[i] < 0) continue;
        auto it = truth_id.emplace(true_assignment[i], num_true);
        if (it.second) ++num_true;
        truth[i] = it.first->second;
    }
    if (num_true == 0) return SBP_E_ARG;
    // the smaller side indexes the rows of the matching (src/evaluate.cpp:157-167)
    const bool truth_rows = num_true < num_found;
    const int32_t nr = truth_rows ? num_true : num_found, nc = truth_rows ? num_found : num_true;
    std::vector<int64_t> t((size_t)nr * nc, 0);
    int64_t counted = 0;
    for (int32_t i = 0; i < N; ++i) {
        if (truth[i] < 0) continue;
        const int32_t r = truth_rows ? truth[i] : assignment[i], c = truth_rows ? assignment[i] : truth[i];
        t[(size_t)r * nc + c]++;
        ++counted;
    }
    // match rows to columns (maximum weight) on the square padding; column j of the matched table is the
    // column matched to row j, so the matched pairs lie on the diagonal.  The matching is O(n^3): above
    // MAX_MATCH communities the table keeps its order and the accuracy is NaN (the other scores do not
    // depend on the order).
    constexpr int32_t MAX_MATCH = 4096;
    const int32_t n = nc;
    std::vector<int32_t> col_of(n);
    for (int32_t j = 0; j < n; ++j) col_of[j] = j;
    if (n <= MAX_MATCH) {
        std::vector<int64_t> sq((size_t)n * n, 0);
        for (int32_t r = 0; r < nr; ++r)
            for (int32_t c = 0; c < nc; ++c) sq[(size_t)r * n + c] = t[(size_t)r * nc + c];
        col_of = max_weight_assignment(sq, n);
    }
    std::vector<int32_t> m((size_t)nr * nc);
    for (int32_t j = 0; j < nc; ++j)
        for (int32_t r = 0; r < nr; ++r) m[(size_t)r * nc + j] = (int32_t)t[(size_t)r * nc + col_of[j]];
    // rows are the algorithm's communities, columns the true ones (src/evaluate.cpp:189-199)
    int32_t rows = nr, cols = nc;
    if (truth_rows) {
        std::vector<int32_t> tr((size_t)nr * nc);
        for (int32_t r = 0; r < nr; ++r)
            for (int32_t c = 0; c < nc; ++c) tr[(size_t)c * nr + r] = m[(size_t)r * nc + c];
        m.swap(tr);
        std::swap(rows, cols);
    }
    if (table_rows) *table_rows = rows;
    if (table_cols) *table_cols = cols;
    if (table_out) std::copy(m.begin(), m.end(), table_out);
    // the reference passes graph.num_vertices() (src/evaluate.cpp:121)
    (void)counted;
    const int rc = sbp_evaluate_table(N, rows, cols, m.data(), out);
    if (rc == SBP_OK && n > MAX_MATCH) out->accuracy = NAN;
    return rc;
}

}  // extern "C"
