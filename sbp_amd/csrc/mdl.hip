// Description length: MDL = E*h(B^2/E) + N*ln(B) - sum_{nz} M_rc * logf(M_rc / (d_out[r] * d_in[c])).
//
// Replaces entropy::mdl (src/entropy.cpp:597-604) and Blockmodel::log_posterior_probability
// (src/blockmodel/blockmodel.cpp:288-306), which copy every robin_map row into four nnz-long vectors.
// Here it is one streaming pass over the row-major cells: each wave takes a contiguous chunk of 64 cells,
// finds the chunk's first row once and walks forward.  The ratio is formed in FP64 and rounded to float
// before logf, as the reference does; terms are accumulated in 128-bit fixed point, so the result is
// bit-identical to any other evaluation order.  Algorithmic bytes: 12 per cell + 8 per block.
#include <math.h>

#include "device_common.hpp"
#include "internal.hpp"

namespace sbpimpl {
using namespace sbpdev;

constexpr int MDL_BLOCK = 256;

__global__ void __launch_bounds__(MDL_BLOCK) k_log_posterior(int64_t nnz, int32_t B, const int32_t *__restrict__ row_ptr,
                                                             const int2 *__restrict__ cells,
                                                             const int32_t *__restrict__ d_out,
                                                             const int32_t *__restrict__ d_in, Fx *__restrict__ partials,
                                                             int32_t *__restrict__ err) {
    __shared__ Fx wave_part[MDL_BLOCK / WAVE];
    const int lane = lane_id(), wid = threadIdx.x / WAVE;
    const int64_t waves = (int64_t)gridDim.x * (MDL_BLOCK / WAVE);
    Fx acc = fx_zero();
    for (int64_t chunk = blockIdx.x * (int64_t)(MDL_BLOCK / WAVE) + wid; chunk * WAVE < nnz; chunk += waves) {
        int64_t first = chunk * WAVE;
        // row of the chunk's first cell: last r with row_ptr[r] <= first
        int lo = 0, hi = B;
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (row_ptr[mid] <= first) lo = mid;
            else hi = mid - 1;
        }
        int64_t i = first + lane;
        if (i < nnz) {
            int r = lo;
            while (row_ptr[r + 1] <= i) ++r;
            int2 cell = cells[i];
            double value = (double)cell.y;
            double dout = (double)d_out[r], din = (double)d_in[cell.x];
            if (dout == 0.0 || din == 0.0) atomicExch(err, 1);
            double ratio = value / (dout * din);
            double term = value * (double)logf_glibc((float)ratio);
            fx_acc(acc, term);
        }
    }
    acc = wave_sum_fx(acc);
    if (lane == 0) wave_part[wid] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        Fx t = wave_part[0];
        for (int w = 1; w < MDL_BLOCK / WAVE; ++w) fx_add(t, wave_part[w]);
        partials[blockIdx.x] = t;
    }
}

__global__ void __launch_bounds__(MDL_BLOCK) k_sum_partials(int n, const Fx *__restrict__ partials, Fx *__restrict__ out) {
    __shared__ Fx wave_part[MDL_BLOCK / WAVE];
    Fx acc = fx_zero();
    for (int i = threadIdx.x; i < n; i += MDL_BLOCK) fx_add(acc, partials[i]);
    acc = wave_sum_fx(acc);
    const int lane = lane_id(), wid = threadIdx.x / WAVE;
    if (lane == 0) wave_part[wid] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        Fx t = wave_part[0];
        for (int w = 1; w < MDL_BLOCK / WAVE; ++w) fx_add(t, wave_part[w]);
        *out = t;
    }
}

double log_posterior(sbp_ctx ctx, sbp_bm bm) {
    hipStream_t st = ctx->stream;
    if (bm->nnz == 0) return 0.0;
    int64_t chunks = (bm->nnz + WAVE - 1) / WAVE;
    int64_t blocks = (chunks + (MDL_BLOCK / WAVE) - 1) / (MDL_BLOCK / WAVE);
    if (blocks > 4096) blocks = 4096;
    Fx *partials = ctx->scratch[S_PARTIAL].as<Fx>(blocks + 2);
    int32_t *err = reinterpret_cast<int32_t *>(partials + blocks + 1);
    SBP_HIP(hipMemsetAsync(err, 0, sizeof(int32_t), st));
    k_log_posterior<<<(int)blocks, MDL_BLOCK, 0, st>>>(bm->nnz, bm->B, bm->row_ptr, bm->row_cells, bm->d_out,
                                                         bm->d_in, partials, err);
    k_sum_partials<<<1, MDL_BLOCK, 0, st>>>((int)blocks, partials, partials + blocks);
    SBP_HIP(hipGetLastError());
    struct {
        Fx total;
        int32_t err;
    } h;
    SBP_HIP(hipMemcpyAsync(&h.total, partials + blocks, sizeof(Fx), hipMemcpyDeviceToHost, st));
    SBP_HIP(hipMemcpyAsync(&h.err, err, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    SBP_HIP(hipStreamSynchronize(st));
    // The reference exits when a non-zero cell has a zero degree (blockmodel.cpp:297-302).
    SBP_CHECK(h.err == 0, SBP_E_STATE, "log_posterior: non-zero cell with zero block degree");
    return (double)h.total.hi + (double)h.total.lo * 0x1p-64;
}

// E*h(x) + N*ln(B), x = B^2/E, h(x) = (1+x)ln(1+x) - x ln x (src/entropy.cpp:600-603), host glibc.
double model_term(int64_t E, int32_t N, int32_t B) {
    double x = pow((double)B, 2) / (double)E;
    double h = ((1 + x) * log(1 + x)) - (x * log(x));
    return ((double)E * h) + ((double)N * log((double)B));
}

__global__ void k_debug_logf(uint32_t first, uint32_t count, float *out) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x)
        out[i] = logf_glibc(__uint_as_float(first + i));
}

void debug_logf(sbp_ctx ctx, uint32_t first, uint32_t count, float *d_out) {
    int blocks = (int)((count + 255) / 256);
    if (blocks > 65536) blocks = 65536;
    if (blocks < 1) blocks = 1;
    k_debug_logf<<<blocks, 256, 0, ctx->stream>>>(first, count, d_out);
    SBP_HIP(hipGetLastError());
}

}  // namespace sbpimpl
