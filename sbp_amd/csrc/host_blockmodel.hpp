// Host copy of a device blockmodel, used to check a device blockmodel against a rebuild from its
// assignment (sbp_bm_validate).  Rows and columns are (index, count) lists in ascending index order.
#pragma once
#include <string.h>

#include <algorithm>
#include <vector>

#include "internal.hpp"

namespace sbpimpl {

struct HostBlockmodel {
    int32_t B = 0;
    std::vector<std::vector<std::pair<int32_t, int64_t>>> row, col;
    std::vector<int64_t> d_out, d_in, d_tot;
    std::vector<int32_t> sizes;

    void download(sbp_ctx ctx, sbp_bm bm) {
        SBP_HIP(hipStreamSynchronize(ctx->stream));
        B = bm->B;
        std::vector<int32_t> rp(B + 1), cp(B + 1), a(B), b(B), c(B), sz(B);
        std::vector<int2> rc(bm->nnz), cc(bm->nnz);
        SBP_HIP(dev_copy(rp.data(), bm->row_ptr, sizeof(int32_t) * (B + 1), hipMemcpyDeviceToHost));
        SBP_HIP(dev_copy(cp.data(), bm->col_ptr, sizeof(int32_t) * (B + 1), hipMemcpyDeviceToHost));
        if (bm->nnz) {
            SBP_HIP(dev_copy(rc.data(), bm->row_cells, sizeof(int2) * bm->nnz, hipMemcpyDeviceToHost));
            SBP_HIP(dev_copy(cc.data(), bm->col_cells, sizeof(int2) * bm->nnz, hipMemcpyDeviceToHost));
        }
        SBP_HIP(dev_copy(a.data(), bm->d_out, sizeof(int32_t) * B, hipMemcpyDeviceToHost));
        SBP_HIP(dev_copy(b.data(), bm->d_in, sizeof(int32_t) * B, hipMemcpyDeviceToHost));
        SBP_HIP(dev_copy(c.data(), bm->d_tot, sizeof(int32_t) * B, hipMemcpyDeviceToHost));
        SBP_HIP(dev_copy(sz.data(), bm->sizes, sizeof(int32_t) * B, hipMemcpyDeviceToHost));
        row.assign(B, {});
        col.assign(B, {});
        for (int32_t r = 0; r < B; ++r) {
            row[r].reserve(rp[r + 1] - rp[r]);
            for (int32_t i = rp[r]; i < rp[r + 1]; ++i) row[r].push_back({rc[i].x, rc[i].y});
            col[r].reserve(cp[r + 1] - cp[r]);
            for (int32_t i = cp[r]; i < cp[r + 1]; ++i) col[r].push_back({cc[i].x, cc[i].y});
        }
        d_out.assign(a.begin(), a.end());
        d_in.assign(b.begin(), b.end());
        d_tot.assign(c.begin(), c.end());
        sizes = sz;
    }

    bool equals(const HostBlockmodel &o) const {
        return B == o.B && row == o.row && col == o.col && d_out == o.d_out && d_in == o.d_in && d_tot == o.d_tot &&
               sizes == o.sizes;
    }
};

}  // namespace sbpimpl
