// Dense path: Omega as an n x n FP64 matrix (n = 3N), inverted by unpivoted
// Gauss-Jordan, dx = Omega^-1 xi.
//
// Reference algorithm: inv (src/cpu_pcg_fp64/MatrixOps.c:29-91) and its GPU
// twin inv_cuda (src/gpu_gj_fp64/inv.cu:221-354): for each pivot i,
// ratio_j = T[j,i]/T[i,i]; T[j,:] -= ratio_j T[i,:]; R[j,:] -= ratio_j R[i,:]
// for j != i; finally R[j,:] /= T[j,j]; a zero pivot gives the zero matrix.
//
// Here T and R share one n x n buffer (in-place Gauss-Jordan): at pivot i the
// columns k > i still hold T (only they are read later), the columns k <= i
// hold R (R[i,k] = 0 for k > i, so only they change).  Column i switches from
// T to R as R[j,i] = -ratio_j, R[i,i] = 1.  That is 2 n^3 flops instead of the
// reference's 4 n^3 and one launch per pivot instead of two; T[i,i] is kept
// in diag[i] for the final row scaling.  The pivot column is double-buffered
// (colbuf) so no thread reads a value another thread of the same step writes.
#include "device_common.h"
#include "gslam_internal.h"

#include <algorithm>
#include <mutex>
#include <string>

namespace gslam {

constexpr int kGjTile = 16;  // 16 x 16 threads per workgroup, one element each

__global__ void __launch_bounds__(kGjTile * kGjTile)
gj_pivot_kernel(double *__restrict__ m, const double *__restrict__ col_in,
                double *__restrict__ col_out, double *__restrict__ diag, int64_t n, int64_t i,
                int32_t *singular)
{
    const int64_t k = blockIdx.x * static_cast<int64_t>(kGjTile) + threadIdx.x;
    const int64_t j = blockIdx.y * static_cast<int64_t>(kGjTile) + threadIdx.y;
    if (j >= n || k >= n) return;
    if (*singular) return;
    const double piv = col_in[i];
    if (piv == 0.0) {  // MatrixOps.c:60-68
        if (j == 0 && k == 0) *singular = 1;
        return;
    }
    double v;
    if (j == i) {
        v = (k == i) ? 1.0 : m[j * n + k];
        if (k == i) {
            m[j * n + k] = 1.0;
            diag[i] = piv;
        }
    } else {
        const double ratio = col_in[j] / piv;
        if (k == i) {
            v = -ratio;
        } else {
            v = m[j * n + k] - ratio * m[i * n + k];
        }
        m[j * n + k] = v;
    }
    if (k == i + 1) col_out[j] = v;
}

__global__ void __launch_bounds__(kThreads)
gj_column_kernel(const double *__restrict__ m, double *__restrict__ col, int64_t n, int64_t c)
{
    const int64_t j = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (j < n) col[j] = m[j * n + c];
}

__global__ void gj_flag_clear_kernel(int32_t *flag) { *flag = 0; }

__global__ void __launch_bounds__(kThreads)
copy64_kernel(const double *__restrict__ src, double *__restrict__ dst, int64_t m)
{
    for (int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; t < m;
         t += static_cast<int64_t>(gridDim.x) * blockDim.x)
        dst[t] = src[t];
}

__global__ void __launch_bounds__(kThreads)
gj_scale_kernel(double *__restrict__ m, const double *__restrict__ diag, int64_t n,
                const int32_t *singular)
{
    const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (t >= n * n) return;
    m[t] = *singular ? 0.0 : m[t] / diag[t / n];
}

// Dense matrix from the BSR blocks (zeros elsewhere).
__global__ void __launch_bounds__(kThreads)
bsr_to_dense_kernel(const int32_t *__restrict__ brow_ptr, const int32_t *__restrict__ bcol,
                    const double *__restrict__ bval, int64_t n_blocks, int32_t n_rows,
                    double *__restrict__ m)
{
    const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (t >= 9 * n_blocks) return;
    const int64_t b = t / 9;
    const int k = static_cast<int>(t - 9 * b);
    // block row by binary search over brow_ptr
    int32_t lo = 0, hi = n_rows - 1;
    while (lo < hi) {
        const int32_t mid = (lo + hi + 1) >> 1;
        if (brow_ptr[mid] <= b) lo = mid; else hi = mid - 1;
    }
    const int64_t dim = 3 * static_cast<int64_t>(n_rows);
    m[(3 * static_cast<int64_t>(lo) + k / 3) * dim + 3 * static_cast<int64_t>(bcol[b]) + k % 3] = bval[t];
}

__global__ void __launch_bounds__(kThreads)
fill_kernel(double *__restrict__ p, int64_t n, double v)
{
    const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (t < n) p[t] = v;
}

// y = M x, one wavefront per row (dot order of MatrixOps.c:123-147 is not
// kept: lanes stride the row, then a butterfly).
__global__ void __launch_bounds__(kThreads)
dense_matvec_kernel(const double *__restrict__ m, const double *__restrict__ x, int64_t n,
                    double *__restrict__ y)
{
    const int64_t row = (blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x) / kWave;
    const int lane = threadIdx.x % kWave;
    if (row >= n) return;
    double s = 0.0;
    for (int64_t k = lane; k < n; k += kWave) s += m[row * n + k] * x[k];
    s = wave_sum(s);
    if (lane == 0) y[row] = s;
}

static inline unsigned int nb1(int64_t n) { return static_cast<unsigned int>((n + kThreads - 1) / kThreads); }

// In-place FP64 inverse of the n x n device matrix m.  work: 2n + n doubles
// (two column buffers and the diagonal) + one int flag.
int gj_invert_fp64_device(hipStream_t s, double *m, int64_t n, double *work, int32_t *flag,
                          int32_t *singular_out)
{
    double *col0 = work, *col1 = work + n, *diag = work + 2 * n;
    hipLaunchKernelGGL(gj_flag_clear_kernel, dim3(1), dim3(1), 0, s, flag);
    hipLaunchKernelGGL(gj_column_kernel, dim3(nb1(n)), dim3(kThreads), 0, s, m, col0, n, int64_t(0));
    const unsigned int gt = static_cast<unsigned int>((n + kGjTile - 1) / kGjTile);
    for (int64_t i = 0; i < n; ++i) {
        double *cin = (i & 1) ? col1 : col0, *cout = (i & 1) ? col0 : col1;
        hipLaunchKernelGGL(gj_pivot_kernel, dim3(gt, gt), dim3(kGjTile, kGjTile), 0, s, m, cin, cout,
                           diag, n, i, flag);
    }
    hipLaunchKernelGGL(gj_scale_kernel, dim3(nb1(n * n)), dim3(kThreads), 0, s, m, diag, n, flag);
    GSLAM_HIP(hipGetLastError());
    GSLAM_HIP(hipMemcpyAsync(singular_out, flag, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    GSLAM_HIP(hipStreamSynchronize(s));
    return GSLAM_OK;
}

int dense_solve(DeviceGraph &g, const gslam_solve_opts &o, const double *b, double *x,
                gslam_solve_stats &st)
{
    const int64_t n = 3 * static_cast<int64_t>(g.h.n);
    if (o.solver == GSLAM_SOLVER_GJ_MIXED) {
        if (n > 131072) {
            set_error("GJ_MIXED limited to 3N <= 131072 (68.7 GB FP32 matrix)");
            return GSLAM_ERR_TOO_LARGE;
        }
        return dense_solve_mixed(g, o, b, x, st);
    }
    if (o.solver == GSLAM_SOLVER_GJ_BLOCKED64) {
        if (n > 92160) {
            set_error("GJ_BLOCKED64 limited to 3N <= 92160 (68 GB FP64 matrix)");
            return GSLAM_ERR_TOO_LARGE;
        }
        return dense_solve_blocked64(g, o, b, x, st);
    }
    if (n > 60000) {
        set_error("dense solve limited to 3N <= 60000 (28.8 GB FP64 matrix)");
        return GSLAM_ERR_TOO_LARGE;
    }
    // the GJ_BLOCKED64 work space, kept on the handle between solves: its
    // padded matrix (np >= n) holds the n x n matrix, its column panel
    // (128 np >= 3n + 8 doubles) the pivot columns and the diagonal
    GSLAM_TRY(g.dense64.reserve(n));
    double *m = g.dense64.a, *work = g.dense64.ct;
    int32_t *flag = g.dense64.flag;
    hipStream_t s = g.stream;
    (void)hipEventRecord(g.ev0, s);
    hipLaunchKernelGGL(fill_kernel, dim3(nb1(n * n)), dim3(kThreads), 0, s, m, n * n, 0.0);
    hipLaunchKernelGGL(bsr_to_dense_kernel, dim3(nb1(9 * g.h.n_blocks)), dim3(kThreads), 0, s,
                       g.brow_ptr, g.bcol, g.bval, g.h.n_blocks, g.h.n, m);
    int32_t singular = 0;
    int rc = gj_invert_fp64_device(s, m, n, work, flag, &singular);
    if (rc == GSLAM_OK) {
        hipLaunchKernelGGL(dense_matvec_kernel, dim3(nb1(n * kWave)), dim3(kThreads), 0, s, m, b,
                           n, x);
        (void)hipEventRecord(g.ev1, s);
        rc = (hipStreamSynchronize(s) == hipSuccess) ? GSLAM_OK : GSLAM_ERR_HIP;
    }
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, g.ev0, g.ev1);
    if (rc != GSLAM_OK) return rc;
    st.ms = ms;
    st.ms_factor = ms;
    st.iters = 0;
    st.converged = singular ? 0 : 1;
    GSLAM_TRY(true_rel_residual(g, b, x, &st.true_rel_res));
    st.rel_res = st.true_rel_res;
    if (singular) {
        set_error("Gauss-Jordan: zero pivot (the solution is the zero vector)");
        return GSLAM_ERR_SINGULAR;
    }
    return GSLAM_OK;
}

// ---- allocation-free entry point (gslam_gj_invert_device) ----
// Scratch of the in-place FP64 Gauss-Jordan -- the two pivot-column buffers,
// the diagonal (3n doubles) and the zero-pivot flag -- in a file-scope
// device array, so that the entry point calls no allocator.  Every GPU the
// library runs on holds its own copy until the process ends.  Host code
// must not pass the array's name to a kernel (it is a host-side stand-in):
// the device address comes from hipGetSymbolAddress, once per device.
constexpr int32_t kNoAllocMaxN = 16384;
constexpr int64_t kScratchDoubles = 3 * static_cast<int64_t>(kNoAllocMaxN) + 8;
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double *g_scratch_dev[kMaxDevices] = {};

static int scratch_address(double **out)
{
    int dev = 0;
    GSLAM_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("gslam_gj_invert_device: device ordinal beyond 64");
        return GSLAM_ERR_ARG;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error(std::string("gslam_gj_invert_device: hipGetSymbolAddress(scratch): ") +
                      hipGetErrorString(e));
            return GSLAM_ERR_HIP;
        }
        g_scratch_dev[dev] = static_cast<double *>(p);
    }
    *out = g_scratch_dev[dev];
    return GSLAM_OK;
}

}  // namespace gslam

extern "C" int gslam_gj_invert_device(const double *a, double *inv, int32_t n)
{
    using namespace gslam;
    if (!a || !inv || n <= 0) return GSLAM_ERR_ARG;
    if (n > kNoAllocMaxN) {
        set_error("gslam_gj_invert_device: n = " + std::to_string(n) + " beyond the " +
                  std::to_string(kNoAllocMaxN) + " its static scratch holds (3n + 8 doubles); "
                  "use gslam_gj_invert");
        return GSLAM_ERR_TOO_LARGE;
    }
    double *work = nullptr;
    GSLAM_TRY(scratch_address(&work));
    int32_t *flag = reinterpret_cast<int32_t *>(work + 3 * static_cast<int64_t>(n));
    const int64_t nn = static_cast<int64_t>(n) * n;
    hipStream_t s = nullptr;  // the null stream: creating one would allocate
    if (a != inv) {
        const unsigned int gr = static_cast<unsigned int>(std::min<int64_t>(4096, nb1(nn)));
        hipLaunchKernelGGL(copy64_kernel, dim3(gr), dim3(kThreads), 0, s, a, inv, nn);
    }
    int32_t singular = 0;
    GSLAM_TRY(gj_invert_fp64_device(s, inv, n, work, flag, &singular));
    if (singular) {
        set_error("Gauss-Jordan: zero pivot (the inverse is the zero matrix)");
        return GSLAM_ERR_SINGULAR;
    }
    return GSLAM_OK;
}

extern "C" int gslam_gj_invert(const double *a, double *inv, int32_t n, int32_t solver)
{
    using namespace gslam;
    if (!a || !inv || n <= 0) return GSLAM_ERR_ARG;
    if (solver == GSLAM_SOLVER_GJ_MIXED) return bgj_invert_host(a, inv, n);
    if (solver == GSLAM_SOLVER_GJ_BLOCKED64) return bgj64_invert_host(a, inv, n);
    if (solver != GSLAM_SOLVER_GJ_FP64) {
        set_error("gslam_gj_invert: solver must be GJ_FP64, GJ_BLOCKED64 or GJ_MIXED");
        return GSLAM_ERR_ARG;
    }
    const int64_t nn = static_cast<int64_t>(n) * n;
    hipStream_t s = nullptr;
    GSLAM_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    double *m = nullptr, *work = nullptr;
    int32_t *flag = nullptr;
    int rc = GSLAM_OK;
    int32_t singular = 0;
    if (hipMalloc(reinterpret_cast<void **>(&m), static_cast<size_t>(nn) * 8) != hipSuccess ||
        hipMalloc(reinterpret_cast<void **>(&work), static_cast<size_t>(3 * n + 8) * 8) != hipSuccess ||
        hipMalloc(reinterpret_cast<void **>(&flag), 64) != hipSuccess) {
        set_error("gslam_gj_invert: out of device memory");
        rc = GSLAM_ERR_NOMEM;
    }
    if (rc == GSLAM_OK && hipMemcpyAsync(m, a, static_cast<size_t>(nn) * 8, hipMemcpyHostToDevice, s) != hipSuccess)
        rc = GSLAM_ERR_HIP;
    if (rc == GSLAM_OK) rc = gj_invert_fp64_device(s, m, n, work, flag, &singular);
    if (rc == GSLAM_OK && (hipMemcpyAsync(inv, m, static_cast<size_t>(nn) * 8, hipMemcpyDeviceToHost, s) != hipSuccess ||
                           hipStreamSynchronize(s) != hipSuccess))
        rc = GSLAM_ERR_HIP;
    (void)hipFree(m);
    (void)hipFree(work);
    (void)hipFree(flag);
    (void)hipStreamDestroy(s);
    if (rc == GSLAM_OK && singular) {
        set_error("Gauss-Jordan: zero pivot (the inverse is the zero matrix)");
        return GSLAM_ERR_SINGULAR;
    }
    return rc;
}
