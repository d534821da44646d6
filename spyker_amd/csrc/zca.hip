// ZCA whitening (src/spyker/cpu/zca.cpp:66-158): fit a mean and the
// whitening matrix U diag(1 / sqrt(max(0, l) + eps)) U^T of the sample
// covariance, apply (x - mean) . trans in place, and split into the positive
// and negative parts.
//
// The reference fits and applies in the input's F32 through a host BLAS and
// LAPACK.  Here the inputs and outputs stay F32 (the reference's interface
// requires F32, opers.cpp:317-347) and everything between them runs in FP64
// on the device: the column means, the covariance (a symmetric rank-k update
// over row chunks), the symmetric eigensolver and the products, so the F32
// results are the FP64 values rounded once.  The dense linear algebra is
// ROCm's device BLAS/LAPACK (rocBLAS dsyrk/dgemm, rocSOLVER dsyevd), opened
// on first use; the data movement around it is the kernels below.
#include <dlfcn.h>

#include <map>
#include <mutex>

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include "common.hpp"

namespace spk
{
namespace
{
constexpr int64_t ZCA_MAX_COLS = 16384;        // one FP64 [cols][cols] matrix is at most 2 GiB
constexpr int64_t ZCA_CHUNK_ELEMS = 1ll << 25;  // FP64 row-chunk buffers of at most 256 MiB

// The device BLAS/LAPACK entry points, resolved from the process's rocBLAS and
// rocSOLVER (the ones already loaded, e.g. by torch, or the system's).
struct LinAlg
{
    decltype(&rocblas_create_handle) create_handle = nullptr;
    decltype(&rocblas_set_stream) set_stream = nullptr;
    decltype(&rocblas_set_atomics_mode) set_atomics_mode = nullptr;
    decltype(&rocblas_dsyrk) dsyrk = nullptr;
    decltype(&rocblas_dgemm) dgemm = nullptr;
    decltype(&rocsolver_dsyevd) dsyevd = nullptr;
    std::map<int, rocblas_handle> handles;
    std::mutex mutex;
    std::string error;
};

template <typename F>
bool resolve(void *lib, const char *name, F &f)
{
    f = reinterpret_cast<F>(dlsym(lib, name));
    return f != nullptr;
}

LinAlg &linalg()
{
    static LinAlg *L = [] {
        auto *l = new LinAlg;
        void *blas = dlopen("librocblas.so.5", RTLD_NOW | RTLD_GLOBAL);
        void *solver = blas ? dlopen("librocsolver.so.0", RTLD_NOW | RTLD_GLOBAL) : nullptr;
        if (!blas || !solver)
        {
            const char *e = dlerror();
            l->error = e ? e : "library not found";
            return l;
        }
        bool ok = resolve(blas, "rocblas_create_handle", l->create_handle) &&
                  resolve(blas, "rocblas_set_stream", l->set_stream) &&
                  resolve(blas, "rocblas_set_atomics_mode", l->set_atomics_mode) &&
                  resolve(blas, "rocblas_dsyrk", l->dsyrk) &&
                  resolve(blas, "rocblas_dgemm", l->dgemm) && resolve(solver, "rocsolver_dsyevd", l->dsyevd);
        if (!ok) l->error = "missing rocBLAS/rocSOLVER entry points";
        return l;
    }();
    SPK_REQUIRE(L->error.empty(), SPK_ERR_UNSUPPORTED,
                "ZCA: rocBLAS/rocSOLVER (librocblas.so.5, librocsolver.so.0) could not be loaded: " << L->error);
    return *L;
}

#define SPK_BLAS(expr)                                                                                         \
    do                                                                                                         \
    {                                                                                                          \
        rocblas_status _b = (expr);                                                                            \
        SPK_REQUIRE(_b == rocblas_status_success, SPK_ERR_RUNTIME, #expr << ": rocBLAS status " << int(_b)); \
    } while (0)

// The handle of the current device bound to `stream`; the caller holds the mutex.
rocblas_handle bound_handle(LinAlg &L, hipStream_t stream)
{
    int dev = 0;
    SPK_HIP(hipGetDevice(&dev));
    auto it = L.handles.find(dev);
    if (it == L.handles.end())
    {
        rocblas_handle h = nullptr;
        SPK_BLAS(L.create_handle(&h));
        // no atomics: every product sums in one order, so repeated calls agree bit for bit
        SPK_BLAS(L.set_atomics_mode(h, rocblas_atomics_not_allowed));
        it = L.handles.emplace(dev, h).first;
    }
    SPK_BLAS(L.set_stream(it->second, stream));
    return it->second;
}

// Column sums of rows [c * per, (c + 1) * per) for chunk c = blockIdx.y; a
// thread owns a column, so each row's loads are coalesced across the block.
__global__ void column_partial(const float *__restrict__ in, double *__restrict__ part, int64_t rows, int64_t cols,
                               int64_t per)
{
    const int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= cols) return;
    const int64_t r0 = int64_t(blockIdx.y) * per, r1 = r0 + per < rows ? r0 + per : rows;
    double s = 0.0;
    for (int64_t r = r0; r < r1; ++r) s += double(in[r * cols + j]);
    part[int64_t(blockIdx.y) * cols + j] = s;
}

// The chunk sums added in chunk order (deterministic), divided by the row count.
__global__ void column_mean(const double *__restrict__ part, double *__restrict__ mean64, float *__restrict__ mean,
                            int64_t chunks, int64_t cols, int64_t rows)
{
    for (int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; j < cols; j += int64_t(gridDim.x) * blockDim.x)
    {
        double s = 0.0;
        for (int64_t c = 0; c < chunks; ++c) s += part[c * cols + j];
        const double m = s / double(rows);
        mean64[j] = m;
        mean[j] = float(m);
    }
}

// mid = in - mean over `n` rows, widened to FP64.
template <typename M>
__global__ void center_rows(const float *__restrict__ in, const M *__restrict__ mean, double *__restrict__ mid,
                            int64_t n, int64_t cols)
{
    const int64_t total = n * cols;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total; i += int64_t(gridDim.x) * blockDim.x)
        mid[i] = double(in[i]) - double(mean[i % cols]);
}

// w[:, k] = v[:, k] / sqrt(max(0, l_k) + eps) (zca.cpp:59-64 and 91-93),
// column-major [d][d].
__global__ void scale_vectors(const double *__restrict__ vec, const double *__restrict__ lam, double *__restrict__ w,
                              int64_t d, double eps)
{
    const int64_t total = d * d;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total; i += int64_t(gridDim.x) * blockDim.x)
    {
        const double l = lam[i / d];
        w[i] = vec[i] * (1.0 / sqrt((0.0 < l ? l : 0.0) + eps));
    }
}

// F32 row-major transform from the FP64 product's upper triangle (the matrix
// is symmetric; both triangles of the output carry the same value).
__global__ void store_symmetric(const double *__restrict__ t64, float *__restrict__ trans, int64_t d)
{
    const int64_t total = d * d;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total; i += int64_t(gridDim.x) * blockDim.x)
    {
        const int64_t r = i / d, c = i % d, u = r < c ? r : c, v = r < c ? c : r;
        trans[i] = float(t64[v * d + u]);
    }
}

// FP64 symmetric matrix from the row-major upper triangle of the F32
// transform, the triangle the reference's product reads (rusym, zca.cpp:116-118).
__global__ void load_symmetric(const float *__restrict__ trans, double *__restrict__ t64, int64_t d)
{
    const int64_t total = d * d;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total; i += int64_t(gridDim.x) * blockDim.x)
    {
        const int64_t r = i / d, c = i % d, u = r < c ? r : c, v = r < c ? c : r;
        t64[i] = double(trans[u * d + v]);
    }
}

__global__ void narrow_rows(const double *__restrict__ y, float *__restrict__ out, int64_t total)
{
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total; i += int64_t(gridDim.x) * blockDim.x)
        out[i] = float(y[i]);
}

// zca.cpp:141-156: [rows][cols] -> [rows][2][cols] as std::max(T(x), T(0))
// and std::max(T(-x), T(0)), with C++'s conversions for the integer types.
template <typename T>
__global__ void split_kernel(const T *__restrict__ in, T *__restrict__ out, int64_t rows, int64_t cols)
{
    const int64_t total = rows * cols;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total; i += int64_t(gridDim.x) * blockDim.x)
    {
        const int64_t r = i / cols, j = i % cols;
        const T a = T(in[i]), b = T(-in[i]), z = T(0);
        out[(r * 2) * cols + j] = a < z ? z : a;
        out[(r * 2 + 1) * cols + j] = b < z ? z : b;
    }
}

int64_t chunk_rows(int64_t rows, int64_t cols)
{
    int64_t c = ZCA_CHUNK_ELEMS / cols;
    if (c < 1) c = 1;
    return c < rows ? c : rows;
}

// in[rows][cols] <- (in - mean) . T, T the symmetric matrix of trans's upper triangle.
void apply(LinAlg &L, rocblas_handle h, float *in, const float *mean, const float *trans, int64_t rows, int64_t cols,
           hipStream_t s)
{
    const int64_t d = cols, crow = chunk_rows(rows, cols);
    Scratch t64(d * d * int64_t(sizeof(double)), s), mid(crow * d * int64_t(sizeof(double)), s),
        y(crow * d * int64_t(sizeof(double)), s);
    load_symmetric<<<flat_grid(d * d, 256), 256, 0, s>>>(trans, t64.as<double>(), d);
    SPK_LAUNCHED();
    const double one = 1.0, zero = 0.0;
    for (int64_t r0 = 0; r0 < rows; r0 += crow)
    {
        const int64_t c = rows - r0 < crow ? rows - r0 : crow;
        float *x = in + r0 * d;
        center_rows<float><<<flat_grid(c * d, 256), 256, 0, s>>>(x, mean, mid.as<double>(), c, d);
        SPK_LAUNCHED();
        // column-major: y (d x c) = T (d x d) . mid (d x c), i.e. row-major y = mid . T
        SPK_BLAS(L.dgemm(h, rocblas_operation_none, rocblas_operation_none, rocblas_int(d), rocblas_int(c),
                         rocblas_int(d), &one, t64.as<double>(), rocblas_int(d), mid.as<double>(), rocblas_int(d),
                         &zero, y.as<double>(), rocblas_int(d)));
        narrow_rows<<<flat_grid(c * d, 256), 256, 0, s>>>(y.as<double>(), x, c * d);
        SPK_LAUNCHED();
    }
}

template <typename F>
void with_split_type(int32_t type, F &&f)
{
    switch (type)
    {
    case SPK_I8: return f(int8_t());
    case SPK_I16: return f(int16_t());
    case SPK_I32: return f(int32_t());
    case SPK_I64: return f(int64_t());
    case SPK_U8: return f(uint8_t());
    case SPK_U16: return f(uint16_t());
    case SPK_U32: return f(uint32_t());
    case SPK_U64: return f(uint64_t());
    case SPK_F16: return f(_Float16());
    case SPK_F32: return f(float());
    case SPK_F64: return f(double());
    default: SPK_REQUIRE(false, SPK_ERR_UNSUPPORTED, "ZCA: data type " << type_name(type) << " is not supported.");
    }
}
}  // namespace
}  // namespace spk

using namespace spk;

extern "C" {

int32_t spk_zca_fit(void *in, void *mean, void *trans, int64_t rows, int64_t cols, double epsilon, int32_t transform,
                    void *stream)
{
    return guard([&] {
        SPK_ARG(rows >= 2, "ZCA: fitting needs at least 2 samples.");
        SPK_ARG(cols >= 1, "ZCA: input has no features.");
        SPK_ARG(cols <= ZCA_MAX_COLS, "ZCA: at most " << ZCA_MAX_COLS << " features are supported.");
        SPK_ARG(rows <= INT32_MAX / 2, "ZCA: too many samples.");
        LinAlg &L = linalg();
        std::lock_guard<std::mutex> lock(L.mutex);
        hipStream_t s = S(stream);
        rocblas_handle h = bound_handle(L, s);
        const int64_t d = cols;

        // column means: chunked partial sums, then the chunks in order
        const int64_t cblocks = cdiv(d, 256);
        int64_t chunks = 2048 / cblocks;
        if (chunks < 1) chunks = 1;
        if (chunks > rows) chunks = rows;
        const int64_t per = cdiv(rows, chunks);
        chunks = cdiv(rows, per);
        Scratch part(chunks * d * int64_t(sizeof(double)), s), mean64(d * int64_t(sizeof(double)), s);
        column_partial<<<dim3(unsigned(cblocks), unsigned(chunks)), 256, 0, s>>>((const float *)in,
                                                                               part.as<double>(), rows, d, per);
        SPK_LAUNCHED();
        column_mean<<<flat_grid(d, 256), 256, 0, s>>>(part.as<double>(), mean64.as<double>(), (float *)mean, chunks, d,
                                                     rows);
        SPK_LAUNCHED();

        // covariance (zca.cpp:83-84): sigma = mid^T mid / (rows - 1), upper triangle
        const int64_t crow = chunk_rows(rows, d);
        Scratch sigma(d * d * int64_t(sizeof(double)), s);
        {
            Scratch mid(crow * d * int64_t(sizeof(double)), s);
            const double alpha = 1.0 / double(rows - 1);
            for (int64_t r0 = 0; r0 < rows; r0 += crow)
            {
                const int64_t c = rows - r0 < crow ? rows - r0 : crow;
                const double beta = r0 == 0 ? 0.0 : 1.0;
                center_rows<double><<<flat_grid(c * d, 256), 256, 0, s>>>((const float *)in + r0 * d,
                                                                         mean64.as<double>(), mid.as<double>(), c, d);
                SPK_LAUNCHED();
                // column-major mid is d x c; sigma (d x d) = mid . mid^T
                SPK_BLAS(L.dsyrk(h, rocblas_fill_upper, rocblas_operation_none, rocblas_int(d), rocblas_int(c), &alpha,
                                 mid.as<double>(), rocblas_int(d), &beta, sigma.as<double>(), rocblas_int(d)));
            }
        }

        // eigenvectors (columns of sigma) and eigenvalues (zca.cpp:86-87)
        Scratch lam(d * int64_t(sizeof(double)), s), off(d * int64_t(sizeof(double)), s),
            info(int64_t(sizeof(rocblas_int)), s);
        SPK_BLAS(L.dsyevd(h, rocblas_evect_original, rocblas_fill_upper, rocblas_int(d), sigma.as<double>(),
                          rocblas_int(d), lam.as<double>(), off.as<double>(), info.as<rocblas_int>()));
        rocblas_int status = 0;
        SPK_HIP(hipMemcpyAsync(&status, info.ptr, sizeof(status), hipMemcpyDeviceToHost, s));
        SPK_HIP(hipStreamSynchronize(s));
        SPK_REQUIRE(status == 0, SPK_ERR_RUNTIME, "ZCA: the eigensolver did not converge (info " << status << ").");

        // trans = U diag(1 / sqrt(max(0, l) + eps)) U^T (zca.cpp:89-97), eps as the F32 the reference takes
        {
            Scratch w(d * d * int64_t(sizeof(double)), s), t64(d * d * int64_t(sizeof(double)), s);
            scale_vectors<<<flat_grid(d * d, 256), 256, 0, s>>>(sigma.as<double>(), lam.as<double>(), w.as<double>(),
                                                               d, double(float(epsilon)));
            SPK_LAUNCHED();
            const double one = 1.0, zero = 0.0;
            SPK_BLAS(L.dgemm(h, rocblas_operation_none, rocblas_operation_transpose, rocblas_int(d), rocblas_int(d),
                             rocblas_int(d), &one, w.as<double>(), rocblas_int(d), sigma.as<double>(), rocblas_int(d),
                             &zero, t64.as<double>(), rocblas_int(d)));
            store_symmetric<<<flat_grid(d * d, 256), 256, 0, s>>>(t64.as<double>(), (float *)trans, d);
            SPK_LAUNCHED();
        }

        // zca.cpp:99-102: the input whitened in place, here with the stored F32 mean and transform
        if (transform) apply(L, h, (float *)in, (const float *)mean, (const float *)trans, rows, d, s);
    });
}

int32_t spk_zca_trans(void *in, const void *mean, const void *trans, int64_t rows, int64_t cols, void *stream)
{
    return guard([&] {
        SPK_ARG(rows >= 0 && cols >= 1, "ZCA: input shape is invalid.");
        SPK_ARG(cols <= ZCA_MAX_COLS, "ZCA: at most " << ZCA_MAX_COLS << " features are supported.");
        if (rows == 0) return;
        LinAlg &L = linalg();
        std::lock_guard<std::mutex> lock(L.mutex);
        hipStream_t s = S(stream);
        rocblas_handle h = bound_handle(L, s);
        apply(L, h, (float *)in, (const float *)mean, (const float *)trans, rows, cols, s);
    });
}

int32_t spk_zca_split(const void *in, void *out, int32_t type, int64_t rows, int64_t cols, void *stream)
{
    return guard([&] {
        SPK_ARG(rows >= 0 && cols >= 0, "ZCA: input shape is invalid.");
        hipStream_t s = S(stream);
        with_split_type(type, [&](auto t0) {
            using T = decltype(t0);
            if (rows * cols == 0) return;
            split_kernel<T><<<flat_grid(rows * cols, 256), 256, 0, s>>>((const T *)in, (T *)out, rows, cols);
            SPK_LAUNCHED();
        });
    });
}
}
