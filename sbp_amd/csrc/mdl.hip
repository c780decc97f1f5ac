// Description length: MDL = E*h(B^2/E) + N*ln(B) - sum_{nz} M_rc * logf(M_rc / (d_out[r] * d_in[c])).
//
// Replaces entropy::mdl (src/entropy.cpp:597-604) and Blockmodel::log_posterior_probability
// (src/blockmodel/blockmodel.cpp:288-306), which copy every robin_map row into four nnz-long vectors.
// Here it is one streaming pass over the row-major cells: each wave takes a contiguous chunk of 64 cells,
// finds the chunk's first row once and walks forward.  The ratio is formed in FP64 and rounded to float
// before logf, as the reference does; terms are accumulated in 128-bit fixed point, so the result is
// bit-identical to any other evaluation order.  Algorithmic bytes: 12 per cell + 8 per block.
#include <math.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "device_common.hpp"
#include "internal.hpp"
#include "np_entropy.hpp"

namespace sbpimpl {
using namespace sbpdev;

constexpr int MDL_BLOCK = 256;

// One partial sum and one error flag per workgroup; nothing needs to be cleared before the launch.  With
// nnz_on_device, nnz is row_ptr[B] (a recorded reduction, launched for the largest nnz).
__global__ void __launch_bounds__(MDL_BLOCK) k_log_posterior(int64_t nnz, int nnz_on_device, int32_t B,
                                                             const int32_t *__restrict__ row_ptr,
                                                             const int2 *__restrict__ cells,
                                                             const int32_t *__restrict__ d_out,
                                                             const int32_t *__restrict__ d_in, Fx *__restrict__ partials,
                                                             int32_t *__restrict__ errs) {
    if (nnz_on_device) nnz = row_ptr[B];
    __shared__ Fx wave_part[MDL_BLOCK / WAVE];
    __shared__ int32_t block_err;
    const int lane = lane_id(), wid = threadIdx.x / WAVE;
    const int64_t waves = (int64_t)gridDim.x * (MDL_BLOCK / WAVE);
    if (threadIdx.x == 0) block_err = 0;
    __syncthreads();
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
            if (dout == 0.0 || din == 0.0) block_err = 1;
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
        errs[blockIdx.x] = block_err;
    }
}

// Sums the n partials into *out and ORs the n error flags into *err_out (errs and err_out may be null).
__global__ void __launch_bounds__(MDL_BLOCK) k_sum_partials(int n, const Fx *__restrict__ partials,
                                                            const int32_t *__restrict__ errs, Fx *__restrict__ out,
                                                            int32_t *__restrict__ err_out) {
    __shared__ Fx wave_part[MDL_BLOCK / WAVE];
    __shared__ int32_t wave_err[MDL_BLOCK / WAVE];
    Fx acc = fx_zero();
    int32_t e = 0;
    for (int i = threadIdx.x; i < n; i += MDL_BLOCK) {
        fx_add(acc, partials[i]);
        if (errs) e |= errs[i];
    }
    acc = wave_sum_fx(acc);
    e = wave_sum(e);
    const int lane = lane_id(), wid = threadIdx.x / WAVE;
    if (lane == 0) {
        wave_part[wid] = acc;
        wave_err[wid] = e;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        Fx t = wave_part[0];
        int32_t x = wave_err[0];
        for (int w = 1; w < MDL_BLOCK / WAVE; ++w) {
            fx_add(t, wave_part[w]);
            x |= wave_err[w];
        }
        *out = t;
        if (err_out) *err_out = x;
    }
}

// Workgroups of the log-posterior reduction for nnz cells, at most max_blocks (the kernel strides over the rest).
static int64_t lp_blocks(int64_t nnz, int64_t max_blocks) {
    const int64_t chunks = (nnz + WAVE - 1) / WAVE;
    const int64_t blocks = (chunks + (MDL_BLOCK / WAVE) - 1) / (MDL_BLOCK / WAVE);
    return std::max<int64_t>(1, std::min(blocks, max_blocks));
}

// Launches the reduction with its partials in `work` (blocks + 1 Fx, then blocks + 1 int32 flags) and reads the
// total back through the context's pinned staging area.
static double lp_run(sbp_ctx ctx, sbp_bm bm, int64_t blocks, void *work) {
    hipStream_t st = ctx->stream;
    Fx *partials = static_cast<Fx *>(work);
    int32_t *errs = reinterpret_cast<int32_t *>(partials + blocks + 1);
    k_log_posterior<<<(int)blocks, MDL_BLOCK, 0, st>>>(bm->nnz, 0, bm->B, bm->row_ptr, bm->row_cells, bm->d_out,
                                                         bm->d_in, partials, errs);
    k_sum_partials<<<1, MDL_BLOCK, 0, st>>>((int)blocks, partials, errs, partials + blocks, errs + blocks);
    SBP_HIP(hipGetLastError());
    struct {
        Fx total;
        int32_t err;
    } h;
    read_small(ctx, {{&h.total, partials + blocks, sizeof(Fx)}, {&h.err, errs + blocks, sizeof(int32_t)}});
    // The reference exits when a non-zero cell has a zero degree (blockmodel.cpp:297-302).
    SBP_CHECK(h.err == 0, SBP_E_STATE, "log_posterior: non-zero cell with zero block degree");
    return (double)h.total.hi + (double)h.total.lo * 0x1p-64;
}

static size_t lp_work_bytes(int64_t blocks) { return sizeof(Fx) * (blocks + 1) + sizeof(int32_t) * (blocks + 1); }

double log_posterior(sbp_ctx ctx, sbp_bm bm) {
    if (bm->nnz == 0) return 0.0;
    const int64_t blocks = lp_blocks(bm->nnz, 4096);
    return lp_run(ctx, bm, blocks, ctx->scratch[S_PARTIAL].get(lp_work_bytes(blocks)));
}

void log_posterior_reserve(sbp_ctx ctx, int64_t max_nnz) {
    ctx->scratch[S_PARTIAL].get(lp_work_bytes(lp_blocks(std::max<int64_t>(max_nnz, 1), 4096)));
}

void log_posterior_device(sbp_ctx ctx, sbp_bm bm, int64_t max_nnz, Fx *out_total, int32_t *out_err) {
    hipStream_t st = ctx->stream;
    const int64_t blocks = lp_blocks(std::max<int64_t>(max_nnz, 1), 4096);
    Fx *partials = static_cast<Fx *>(ctx->scratch[S_PARTIAL].get(lp_work_bytes(blocks)));
    int32_t *errs = reinterpret_cast<int32_t *>(partials + blocks + 1);
    k_log_posterior<<<(int)blocks, MDL_BLOCK, 0, st>>>(max_nnz, 1, bm->B, bm->row_ptr, bm->row_cells, bm->d_out,
                                                         bm->d_in, partials, errs);
    k_sum_partials<<<1, MDL_BLOCK, 0, st>>>((int)blocks, partials, errs, out_total, out_err);
    SBP_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ allocation-free MDL (sbp_mdl_noalloc)
//
// The reduction's work area lives in a fixed device array instead of the context's growable scratch, so the call
// allocates nothing.  The grid is capped at NOALLOC_BLOCKS workgroups (1,024 x 4 wavefronts, enough to cover the
// 256 CUs; the kernel strides over any number of cells), so every blockmodel fits.  The array is one per device
// and holds the last call's partials; calls that overlap on one device race on it.
constexpr int64_t NOALLOC_BLOCKS = 1024;
constexpr size_t NOALLOC_BYTES = sizeof(Fx) * (NOALLOC_BLOCKS + 1) + sizeof(int32_t) * (NOALLOC_BLOCKS + 1);
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex noalloc_mu;
static void *noalloc_ptr[NOALLOC_MAX_DEVICES];  // device address of `scratch` on each device, set at set-up

// Looks up the array's device address on `device` (current), once per device; called by sbp_create.  In host code
// the name `scratch` is only a host-side handle for the symbol, so kernels get this address instead.
void noalloc_setup(int device) {
    SBP_CHECK(device >= 0 && device < NOALLOC_MAX_DEVICES, SBP_E_ARG, "device index beyond the supported range");
    std::lock_guard<std::mutex> lk(noalloc_mu);
    if (noalloc_ptr[device]) return;
    void *ptr = nullptr;
    SBP_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    noalloc_ptr[device] = ptr;
}

double log_posterior_noalloc(sbp_ctx ctx, sbp_bm bm) {
    void *work = nullptr;
    {
        std::lock_guard<std::mutex> lk(noalloc_mu);
        work = (ctx->device >= 0 && ctx->device < NOALLOC_MAX_DEVICES) ? noalloc_ptr[ctx->device] : nullptr;
    }
    SBP_CHECK(work != nullptr, SBP_E_STATE, "sbp_mdl_noalloc: the device's work area was not set up (sbp_create)");
    if (bm->nnz == 0) return 0.0;
    const int64_t blocks = lp_blocks(bm->nnz, NOALLOC_BLOCKS);
    SBP_CHECK(lp_work_bytes(blocks) <= NOALLOC_BYTES, SBP_E_ARG,
              "sbp_mdl_noalloc: the reduction needs more than the fixed work area");
    return lp_run(ctx, bm, blocks, work);
}

// E*h(x) + N*ln(B), x = B^2/E, h(x) = (1+x)ln(1+x) - x ln x (src/entropy.cpp:600-603), host glibc.
double model_term(int64_t E, int32_t N, int32_t B) {
    double x = pow((double)B, 2) / (double)E;
    double h = ((1 + x) * log(1 + x)) - (x * log(x));
    return ((double)E * h) + ((double)N * log((double)B));
}

// ------------------------------------------------------------------ nonparametric description length

// fastlgamma / fastlbinom / get_edges_dl on the host with glibc lgamma (include/fastlgamma.hpp:19-27,
// include/entropy.hpp:145-149, src/entropy.cpp:733-738)
double host_lgamma(int64_t x) {
    if (x <= 0) return INFINITY;
    int sign;
    return lgamma_r((double)x, &sign);
}
double host_lbinom(int64_t n, int64_t k) {
    if (n == 0 || k == 0 || k > n) return 0;
    return ((host_lgamma(n + 1) - host_lgamma(k + 1)) - host_lgamma(n - k + 1));
}
double host_edges_dl(int64_t B, int64_t E) { return host_lbinom(B * B + E - 1, E); }

// Device table of fastlgamma(x) for x < max(E, N) + 2 (every cell value + 1 and block size + 1), kept on the
// context and grown when a larger graph needs it.
static const double *lgamma_table(sbp_ctx ctx, int64_t n, int64_t *n_out) {
    if (!ctx->lgamma || ctx->lgamma_n < n) {
        std::vector<double> h((size_t)n);
        for (int64_t x = 0; x < n; ++x) h[(size_t)x] = host_lgamma(x);
        SBP_HIP(hipStreamSynchronize(ctx->stream));
        if (ctx->lgamma) SBP_HIP(dev_free(ctx->lgamma));
        ctx->lgamma = nullptr;
        SBP_HIP(dev_malloc(&ctx->lgamma, sizeof(double) * (size_t)n));
        SBP_HIP(dev_copy(ctx->lgamma, h.data(), sizeof(double) * (size_t)n, hipMemcpyHostToDevice));
        ctx->lgamma_n = n;
    }
    *n_out = ctx->lgamma_n;
    return ctx->lgamma;
}

__global__ void k_count_nonempty(int32_t B, const int32_t *__restrict__ sizes, int32_t *__restrict__ out) {
    int32_t local = 0;
    for (int32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < B; b += gridDim.x * blockDim.x) local += sizes[b] > 0;
    local = wave_sum(local);
    if (lane_id() == 0 && local) atomicAdd(out, local);
}

int64_t count_nonempty(sbp_ctx ctx, sbp_bm bm) {
    hipStream_t st = ctx->stream;
    int32_t *d = ctx->scratch[S_NP].as<int32_t>(2);
    SBP_HIP(hipMemsetAsync(d, 0, sizeof(int32_t), st));
    int blocks = (int)std::min<int64_t>((bm->B + 255) / 256, 1024);
    k_count_nonempty<<<std::max(blocks, 1), 256, 0, st>>>(bm->B, bm->sizes, d);
    SBP_HIP(hipGetLastError());
    int32_t h = 0;
    read_small(ctx, {{&h, d, sizeof(int32_t)}});
    return h;
}

// Constants of one launch: the lgamma table and the block-count terms around the current non-empty count.
NpConst np_const(sbp_ctx ctx, sbp_bm bm) {
    NpConst c{};
    c.dc = bm->dc;
    if (bm->dc) {
        SBP_CHECK(bm->hist_ready, SBP_E_STATE, "degree histograms not built");
        c.hist = bm->hist;
        c.hist_shift = 64 - bm->hist_log2;
        c.hist_mask = (uint32_t)((1ull << bm->hist_log2) - 1);
    }
    c.ln2 = log(2.0);
    c.lnpi = log(M_PI);
    const int64_t N = bm->N, E = bm->E;
    c.lg = lgamma_table(ctx, std::max<int64_t>(E, N) + 2, &c.n);
    const int64_t ne = count_nonempty(ctx, bm);
    for (int i = 0; i < 3; ++i) {
        c.lb[i] = host_lbinom(N - 1, ne - 2 + i);
        c.edl[i] = host_edges_dl(ne - 1 + i, E);
    }
    return c;
}

// sparse_entropy (src/entropy.cpp:623-655): eterm over the cells + vterm over the blocks, fixed point
// and, degree-corrected, get_deg_entropy (src/entropy.cpp:619-623) over the N vertices
__global__ void __launch_bounds__(MDL_BLOCK) k_np_sparse(int64_t nnz, int32_t B, const int2 *__restrict__ cells,
                                                         const int32_t *__restrict__ d_out,
                                                         const int32_t *__restrict__ d_in,
                                                         const int32_t *__restrict__ sizes, int32_t N,
                                                         const int64_t *__restrict__ out_off,
                                                         const int64_t *__restrict__ in_off, NpConst c,
                                                         Fx *__restrict__ partials) {
    __shared__ Fx wave_part[MDL_BLOCK / WAVE];
    const int lane = lane_id(), wid = threadIdx.x / WAVE;
    Fx acc = fx_zero();
    const int64_t stride = (int64_t)gridDim.x * MDL_BLOCK;
    const int64_t total = nnz + B + (c.dc ? N : 0);
    for (int64_t i = blockIdx.x * (int64_t)MDL_BLOCK + threadIdx.x; i < total; i += stride) {
        if (i < nnz) {
            fx_acc(acc, np_eterm(c, cells[i].y));
        } else if (i < nnz + B) {
            const int32_t b = (int32_t)(i - nnz);
            fx_acc(acc, np_vterm(c, d_out[b], d_in[b], sizes[b]));
        } else {
            const int32_t v = (int32_t)(i - nnz - B);
            fx_acc(acc, -np_flg(c, in_off[v + 1] - in_off[v] + 1) - np_flg(c, out_off[v + 1] - out_off[v] + 1));
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

// get_deg_dl_dist (src/entropy.cpp:719-742), directed: per block log_q(d_out, size) + log_q(d_in, size) +
// 2 lgamma(size + 1) (the out-histogram of a block counts each of its vertices once), and -lgamma(count + 1)
// per histogram entry.  Summed in double (a count that has gone to -1 gives -inf, as in the reference).
__global__ void __launch_bounds__(MDL_BLOCK) k_np_deg_dl(int32_t B, const int32_t *__restrict__ d_out,
                                                         const int32_t *__restrict__ d_in,
                                                         const int32_t *__restrict__ sizes, int64_t nslots,
                                                         NpConst c, double *__restrict__ partials) {
    __shared__ double wave_part[MDL_BLOCK / WAVE];
    const int lane = lane_id(), wid = threadIdx.x / WAVE;
    double acc = 0.0;
    const int64_t stride = (int64_t)gridDim.x * MDL_BLOCK;
    for (int64_t i = blockIdx.x * (int64_t)MDL_BLOCK + threadIdx.x; i < B + nslots; i += stride) {
        if (i < B) {
            const int32_t b = (int32_t)i;
            acc += dc_log_q(c, (unsigned long long)d_out[b], (unsigned long long)sizes[b]);
            acc += dc_log_q(c, (unsigned long long)d_in[b], (unsigned long long)sizes[b]);
            acc += 2 * np_flg(c, (long long)sizes[b] + 1);
        } else {
            const HistSlot e = c.hist[i - B];
            if (e.key != HIST_EMPTY && e.val != 0) acc -= np_flg(c, (long long)e.val + 1);
        }
    }
    acc = wave_sum(acc);
    if (lane == 0) wave_part[wid] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < MDL_BLOCK / WAVE; ++w) t += wave_part[w];
        partials[blockIdx.x] = t;
    }
}

static double np_deg_dl(sbp_ctx ctx, sbp_bm bm, const NpConst &c) {
    hipStream_t st = ctx->stream;
    const int64_t nslots = int64_t(1) << bm->hist_log2, total = bm->B + nslots;
    const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>((total + MDL_BLOCK - 1) / MDL_BLOCK, 1024));
    double *partials = reinterpret_cast<double *>(ctx->scratch[S_NP].as<double>((size_t)blocks + 4));
    k_np_deg_dl<<<(int)blocks, MDL_BLOCK, 0, st>>>(bm->B, bm->d_out, bm->d_in, bm->sizes, nslots, c, partials);
    SBP_HIP(hipGetLastError());
    std::vector<double> h((size_t)blocks);
    SBP_HIP(hipMemcpyAsync(h.data(), partials, sizeof(double) * blocks, hipMemcpyDeviceToHost, st));
    SBP_HIP(hipStreamSynchronize(st));
    double S = 0.0;
    for (double x : h) S += x;
    return S;
}

// nonparametric::mdl (src/entropy.cpp:740-767).  parts (optional): sparse entropy, partition dl, edges dl,
// degree dl.
double np_mdl(sbp_ctx ctx, sbp_graph g, sbp_bm bm, double *parts) {
    hipStream_t st = ctx->stream;
    bm_set_model(ctx, g, bm, bm->np, bm->dc);
    const NpConst c = np_const(ctx, bm);
    const int64_t total = bm->nnz + bm->B + (bm->dc ? (int64_t)g->N : 0);
    int64_t blocks = std::min<int64_t>((total + MDL_BLOCK - 1) / MDL_BLOCK, 4096);
    blocks = std::max<int64_t>(blocks, 1);
    Fx *partials = ctx->scratch[S_PARTIAL].as<Fx>(blocks + 2);
    k_np_sparse<<<(int)blocks, MDL_BLOCK, 0, st>>>(bm->nnz, bm->B, bm->row_cells, bm->d_out, bm->d_in, bm->sizes,
                                                     g->N, g->out_off, g->in_off, c, partials);
    k_sum_partials<<<1, MDL_BLOCK, 0, st>>>((int)blocks, partials, nullptr, partials + blocks, nullptr);
    SBP_HIP(hipGetLastError());
    Fx h;
    std::vector<int32_t> sizes(bm->B);
    SBP_HIP(hipMemcpyAsync(&h, partials + blocks, sizeof(Fx), hipMemcpyDeviceToHost, st));
    SBP_HIP(hipMemcpyAsync(sizes.data(), bm->sizes, sizeof(int32_t) * bm->B, hipMemcpyDeviceToHost, st));
    SBP_HIP(hipStreamSynchronize(st));
    const double S = (double)h.hi + (double)h.lo * 0x1p-64;
    // get_partition_dl (src/entropy.cpp:657-669), in block order
    int64_t ne = 0;
    for (int32_t b = 0; b < bm->B; ++b) ne += sizes[b] > 0;
    const int64_t N = g->N;
    double part = 0;
    part += host_lbinom(N - 1, ne - 1);
    part += host_lgamma(N + 1);
    for (int32_t b = 0; b < bm->B; ++b) part -= host_lgamma(sizes[b] + 1);
    part += (double)logf((float)N);  // fastlog(N)
    const double edges = host_edges_dl(ne, g->E);
    const double deg = bm->dc ? np_deg_dl(ctx, bm, c) : 0.0;  // get_deg_dl_dist: 0 without degree correction
    double S_dl = 0;
    S_dl += part;
    S_dl += deg;
    S_dl += edges;
    if (parts) {
        parts[0] = S;
        parts[1] = part;
        parts[2] = edges;
        parts[3] = deg;
    }
    return S + S_dl * 1.0;
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
