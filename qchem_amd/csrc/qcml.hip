// qcml.hip -- the learned initial-guess network: kernels and the C ABI of include/qcml.h.
//
// Kernels (gfx950):
//   k_gemm        C = alpha op(A) op(B) + beta C.  256 threads = 4 wavefronts per 64 x 64 tile of C,
//                 each wavefront a 32 x 32 quadrant as 2 x 2 v_mfma_f32_16x16x16bf16_1k tiles; the
//                 K loop stages 64 x 16 slices of A and B through LDS as BF16 (FP32 operands are
//                 rounded to nearest-even on the way in).  FP32 accumulation.
//   k_bias_act    Z += b (row broadcast); H = SiLU(Z) (or H = Z for the output layer).
//   k_act_bwd     dZ = dH * SiLU'(Z).
//   k_colsum      db = column sums of dZ.
//   k_mse         dY = 2 (Y - T) / (B nout); per-block partial sums of (Y - T)^2.
//   k_adam        Adam on the FP32 master weights, then the BF16 copy.
// MFMA 16x16x16 operand layout (per lane l of the wavefront): A holds A[l % 16][4 (l / 16) + i],
// B holds B[4 (l / 16) + i][l % 16], the accumulator D[4 (l / 16) + i][l % 16], i = 0..3.
#include "qcml.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <new>
#include <vector>

namespace {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short f32_to_bf16(float f)
{
    const unsigned int u = __float_as_uint(f);
    return (unsigned short)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);  // round to nearest even
}

template <bool BF16>
__device__ __forceinline__ unsigned short load_bf16(const void* p, long long idx)
{
    if constexpr (BF16) return static_cast<const unsigned short*>(p)[idx];
    else return f32_to_bf16(static_cast<const float*>(p)[idx]);
}

constexpr int kTile = 64, kK = 16;

template <bool ABF, bool BBF>
__global__ __launch_bounds__(256) void k_gemm(int M, int N, int K, const void* __restrict__ A, long long sam,
                                              long long sak, const void* __restrict__ B, long long sbk, long long sbn,
                                              float* __restrict__ C, long long ldc, float alpha, float beta)
{
    __shared__ unsigned short As[kTile][kK + 2];  // [m][k], padded
    __shared__ unsigned short Bs[kK][kTile + 2];  // [k][n]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m0 = blockIdx.y * kTile, n0 = blockIdx.x * kTile;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    v4f acc[2][2];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) acc[a][b] = v4f{0.f, 0.f, 0.f, 0.f};
    const int r = lane & 15, q = lane >> 4;
    for (int k0 = 0; k0 < K; k0 += kK) {
        __syncthreads();
        for (int e = tid; e < kTile * kK; e += 256) {
            const int mm = e / kK, kk = e % kK;  // A slice
            const int gm = m0 + mm, gk = k0 + kk;
            As[mm][kk] = (gm < M && gk < K) ? load_bf16<ABF>(A, gm * sam + gk * sak) : (unsigned short)0;
            const int kb = e / kTile, nn = e % kTile;  // B slice
            const int gkb = k0 + kb, gn = n0 + nn;
            Bs[kb][nn] = (gkb < K && gn < N) ? load_bf16<BBF>(B, gkb * sbk + gn * sbn) : (unsigned short)0;
        }
        __syncthreads();
        v4s af[2], bf[2];
        for (int t = 0; t < 2; ++t) {
            const int row = wm + t * 16 + r, col = wn + t * 16 + r;
            for (int i = 0; i < 4; ++i) {
                af[t][i] = (short)As[row][4 * q + i];
                bf[t][i] = (short)Bs[4 * q + i][col];
            }
        }
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(af[a], bf[b], acc[a][b], 0, 0, 0);
    }
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int i = 0; i < 4; ++i) {
                const int gm = m0 + wm + a * 16 + 4 * q + i, gn = n0 + wn + b * 16 + r;
                if (gm < M && gn < N) {
                    float* c = &C[gm * ldc + gn];
                    *c = alpha * acc[a][b][i] + (beta != 0.f ? beta * *c : 0.f);
                }
            }
}

__device__ __forceinline__ float silu(float z) { return z / (1.f + __expf(-z)); }

__global__ void k_bias_act(float* __restrict__ Z, const float* __restrict__ b, float* __restrict__ H, int M, int N,
                           int act)
{
    const long long n = (long long)M * N;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const float z = Z[i] + b[i % N];
        Z[i] = z;
        H[i] = act ? silu(z) : z;
    }
}

__global__ void k_act_bwd(const float* __restrict__ Z, float* __restrict__ dH, long long n)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const float z = Z[i], s = 1.f / (1.f + __expf(-z));
        dH[i] *= s * (1.f + z * (1.f - s));
    }
}

__global__ void k_colsum(const float* __restrict__ X, float* __restrict__ out, int M, int N)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    float s = 0.f;
    for (int r = 0; r < M; ++r) s += X[(long long)r * N + c];
    out[c] = s;
}

__global__ void k_mse(const float* __restrict__ Y, const float* __restrict__ T, float* __restrict__ dY, long long n,
                      float scale, float* __restrict__ partial)
{
    __shared__ float red[256];
    float s = 0.f;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const float d = Y[i] - T[i];
        dY[i] = 2.f * scale * d;
        s += d * d;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                       unsigned short* __restrict__ wb, long long n, float lr, float c1, float c2)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = 0.9f * m[i] + 0.1f * gi;
        const float vi = 0.999f * v[i] + 0.001f * gi * gi;
        m[i] = mi;
        v[i] = vi;
        const float wi = w[i] - lr * (mi * c1) / (sqrtf(vi * c2) + 1e-8f);
        w[i] = wi;
        if (wb) wb[i] = f32_to_bf16(wi);
    }
}

__global__ void k_to_bf16(const float* __restrict__ w, unsigned short* __restrict__ wb, long long n)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        wb[i] = f32_to_bf16(w[i]);
}

int grid_for(long long n) { return (int)std::min<long long>((n + 255) / 256, 4096); }

hipError_t gemm(int M, int N, int K, const void* A, bool abf, long long sam, long long sak, const void* B, bool bbf,
                long long sbk, long long sbn, float* C, long long ldc, float alpha, float beta, hipStream_t s)
{
    if (M <= 0 || N <= 0) return hipSuccess;
    const dim3 grid((N + kTile - 1) / kTile, (M + kTile - 1) / kTile), block(256);
    if (abf && bbf) hipLaunchKernelGGL((k_gemm<true, true>), grid, block, 0, s, M, N, K, A, sam, sak, B, sbk, sbn, C, ldc, alpha, beta);
    else if (abf) hipLaunchKernelGGL((k_gemm<true, false>), grid, block, 0, s, M, N, K, A, sam, sak, B, sbk, sbn, C, ldc, alpha, beta);
    else if (bbf) hipLaunchKernelGGL((k_gemm<false, true>), grid, block, 0, s, M, N, K, A, sam, sak, B, sbk, sbn, C, ldc, alpha, beta);
    else hipLaunchKernelGGL((k_gemm<false, false>), grid, block, 0, s, M, N, K, A, sam, sak, B, sbk, sbn, C, ldc, alpha, beta);
    return hipGetLastError();
}

}  // namespace

struct qcml_net {
    int device = 0;
    int nin = 0, width = 0, nout = 0, maxb = 0;
    long long nparam = 0;
    long long off[6] = {};   // W1 b1 W2 b2 W3 b3 within the flat arrays
    long long cnt[6] = {};
    float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;  // flat FP32 master, grad, Adam moments
    unsigned short* wb = nullptr;                                   // flat BF16 copy
    float *Z1 = nullptr, *H1 = nullptr, *Z2 = nullptr, *H2 = nullptr, *Y = nullptr, *dY = nullptr, *dH2 = nullptr,
          *dH1 = nullptr, *partial = nullptr;
    long long step = 0;
};

#define HIPCHK(x)                                          \
    do {                                                   \
        if ((x) != hipSuccess) return 3;                   \
    } while (0)

namespace {

void free_net(qcml_net* n)
{
    if (!n) return;
    hipSetDevice(n->device);
    hipDeviceSynchronize();
    void* p[] = {n->w, n->g, n->m, n->v, n->wb, n->Z1, n->H1, n->Z2, n->H2, n->Y, n->dY, n->dH2, n->dH1, n->partial};
    for (void* x : p)
        if (x) hipFree(x);
    delete n;
}

template <typename T>
hipError_t dalloc(T** p, long long count)
{
    *p = nullptr;
    return hipMalloc((void**)p, (size_t)std::max<long long>(count, 1) * sizeof(T));
}

// Forward pass for B rows: Z1/H1, Z2/H2, Y (device).
int forward(qcml_net* n, const float* X, int B, float* Y, hipStream_t s)
{
    const int H = n->width;
    const unsigned short* W1 = n->wb + n->off[0];
    const unsigned short* W2 = n->wb + n->off[2];
    const unsigned short* W3 = n->wb + n->off[4];
    HIPCHK(gemm(B, H, n->nin, X, false, n->nin, 1, W1, true, H, 1, n->Z1, H, 1.f, 0.f, s));
    hipLaunchKernelGGL(k_bias_act, dim3(grid_for((long long)B * H)), dim3(256), 0, s, n->Z1, n->w + n->off[1], n->H1, B, H, 1);
    HIPCHK(gemm(B, H, H, n->H1, false, H, 1, W2, true, H, 1, n->Z2, H, 1.f, 0.f, s));
    hipLaunchKernelGGL(k_bias_act, dim3(grid_for((long long)B * H)), dim3(256), 0, s, n->Z2, n->w + n->off[3], n->H2, B, H, 1);
    HIPCHK(gemm(B, n->nout, H, n->H2, false, H, 1, W3, true, n->nout, 1, Y, n->nout, 1.f, 0.f, s));
    hipLaunchKernelGGL(k_bias_act, dim3(grid_for((long long)B * n->nout)), dim3(256), 0, s, Y, n->w + n->off[5], Y, B,
                       n->nout, 0);
    HIPCHK(hipGetLastError());
    return 0;
}

// Forward + backward into n->g; per-block loss partials in n->partial (nblk entries).
int backward(qcml_net* n, const float* X, const float* T, int B, hipStream_t s, int* nblk)
{
    const int H = n->width, O = n->nout;
    int rc = forward(n, X, B, n->Y, s);
    if (rc) return rc;
    const long long ny = (long long)B * O;
    *nblk = grid_for(ny);
    hipLaunchKernelGGL(k_mse, dim3(*nblk), dim3(256), 0, s, n->Y, T, n->dY, ny, 1.f / (float)ny, n->partial);
    const unsigned short* W2 = n->wb + n->off[2];
    const unsigned short* W3 = n->wb + n->off[4];
    // layer 3: dW3 = H2^T dY, db3 = colsum dY, dH2 = dY W3^T
    HIPCHK(gemm(H, O, B, n->H2, false, 1, H, n->dY, false, O, 1, n->g + n->off[4], O, 1.f, 0.f, s));
    hipLaunchKernelGGL(k_colsum, dim3((O + 255) / 256), dim3(256), 0, s, n->dY, n->g + n->off[5], B, O);
    HIPCHK(gemm(B, H, O, n->dY, false, O, 1, W3, true, 1, O, n->dH2, H, 1.f, 0.f, s));
    hipLaunchKernelGGL(k_act_bwd, dim3(grid_for((long long)B * H)), dim3(256), 0, s, n->Z2, n->dH2, (long long)B * H);
    // layer 2
    HIPCHK(gemm(H, H, B, n->H1, false, 1, H, n->dH2, false, H, 1, n->g + n->off[2], H, 1.f, 0.f, s));
    hipLaunchKernelGGL(k_colsum, dim3((H + 255) / 256), dim3(256), 0, s, n->dH2, n->g + n->off[3], B, H);
    HIPCHK(gemm(B, H, H, n->dH2, false, H, 1, W2, true, 1, H, n->dH1, H, 1.f, 0.f, s));
    hipLaunchKernelGGL(k_act_bwd, dim3(grid_for((long long)B * H)), dim3(256), 0, s, n->Z1, n->dH1, (long long)B * H);
    // layer 1
    HIPCHK(gemm(n->nin, H, B, X, false, 1, n->nin, n->dH1, false, H, 1, n->g + n->off[0], H, 1.f, 0.f, s));
    hipLaunchKernelGGL(k_colsum, dim3((H + 255) / 256), dim3(256), 0, s, n->dH1, n->g + n->off[1], B, H);
    HIPCHK(hipGetLastError());
    return 0;
}

int read_loss(qcml_net* n, int nblk, hipStream_t s, float* loss, long long ny)
{
    std::vector<float> part(nblk);
    HIPCHK(hipMemcpyAsync(part.data(), n->partial, nblk * sizeof(float), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    double t = 0.0;
    for (float x : part) t += x;
    *loss = (float)(t / (double)ny);
    return 0;
}

}  // namespace

extern "C" {

int qcml_create(int32_t nin, int32_t width, int32_t nout, int32_t max_batch, uint64_t seed, int32_t device,
                qcml_net** out)
{
    if (!out || nin < 1 || width < 1 || nout < 1 || max_batch < 1) return 1;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return 5;
    if (device < 0 || device >= ndev) return 1;
    qcml_net* n = new (std::nothrow) qcml_net();
    if (!n) return 4;
    n->device = device;
    n->nin = nin;
    n->width = width;
    n->nout = nout;
    n->maxb = max_batch;
    const long long c[6] = {(long long)nin * width, width, (long long)width * width, width, (long long)width * nout, nout};
    long long o = 0;
    for (int i = 0; i < 6; ++i) {
        n->off[i] = o;
        n->cnt[i] = c[i];
        o += c[i];
    }
    n->nparam = o;
    auto fail = [&]() { free_net(n); return 3; };
    if (hipSetDevice(device) != hipSuccess) return fail();
    const long long bh = (long long)max_batch * width, bo = (long long)max_batch * nout;
    if (dalloc(&n->w, o) || dalloc(&n->g, o) || dalloc(&n->m, o) || dalloc(&n->v, o) || dalloc(&n->wb, o) ||
        dalloc(&n->Z1, bh) || dalloc(&n->H1, bh) || dalloc(&n->Z2, bh) || dalloc(&n->H2, bh) || dalloc(&n->Y, bo) ||
        dalloc(&n->dY, bo) || dalloc(&n->dH2, bh) || dalloc(&n->dH1, bh) || dalloc(&n->partial, 4096))
        return fail();
    // He-uniform weights from a SplitMix64 stream, zero biases
    std::vector<float> host(o, 0.f);
    uint64_t st = seed;
    auto next = [&]() {
        uint64_t z = (st += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    };
    const int fan_in[3] = {nin, width, width};
    for (int l = 0; l < 3; ++l) {
        const double lim = std::sqrt(6.0 / fan_in[l]) * (l == 2 ? 0.5 : 1.0);
        for (long long i = 0; i < n->cnt[2 * l]; ++i)
            host[n->off[2 * l] + i] = (float)(lim * (2.0 * ((next() >> 11) * (1.0 / 9007199254740992.0)) - 1.0));
    }
    const int rc = qcml_set_params(n, host.data());
    if (rc) {
        free_net(n);
        return rc;
    }
    *out = n;
    return 0;
}

int64_t qcml_nparams(const qcml_net* n) { return n ? n->nparam : -1; }

int qcml_get_params(qcml_net* n, float* host)
{
    if (!n || !host) return 1;
    HIPCHK(hipSetDevice(n->device));
    HIPCHK(hipDeviceSynchronize());  // updates may have run on any caller stream
    HIPCHK(hipMemcpy(host, n->w, n->nparam * sizeof(float), hipMemcpyDeviceToHost));
    return 0;
}

int qcml_set_params(qcml_net* n, const float* host)
{
    if (!n || !host) return 1;
    HIPCHK(hipSetDevice(n->device));
    HIPCHK(hipMemcpy(n->w, host, n->nparam * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(hipMemset(n->m, 0, n->nparam * sizeof(float)));
    HIPCHK(hipMemset(n->v, 0, n->nparam * sizeof(float)));
    hipLaunchKernelGGL(k_to_bf16, dim3(grid_for(n->nparam)), dim3(256), 0, (hipStream_t)0, n->w, n->wb, n->nparam);
    HIPCHK(hipGetLastError());
    HIPCHK(hipDeviceSynchronize());
    n->step = 0;
    return 0;
}

int qcml_forward(qcml_net* n, const float* dX, int32_t B, float* dY, void* stream)
{
    if (!n || !dX || !dY || B < 1 || B > n->maxb) return 1;
    HIPCHK(hipSetDevice(n->device));
    return forward(n, dX, B, dY, (hipStream_t)stream);
}

int qcml_train_step(qcml_net* n, const float* dX, const float* dT, int32_t B, float lr, float* loss, void* stream)
{
    if (!n || !dX || !dT || B < 1 || B > n->maxb) return 1;
    HIPCHK(hipSetDevice(n->device));
    hipStream_t s = (hipStream_t)stream;
    int nblk = 0;
    int rc = backward(n, dX, dT, B, s, &nblk);
    if (rc) return rc;
    n->step += 1;
    const float c1 = (float)(1.0 / (1.0 - std::pow(0.9, (double)n->step)));
    const float c2 = (float)(1.0 / (1.0 - std::pow(0.999, (double)n->step)));
    hipLaunchKernelGGL(k_adam, dim3(grid_for(n->nparam)), dim3(256), 0, s, n->w, n->g, n->m, n->v, n->wb, n->nparam,
                       lr, c1, c2);
    HIPCHK(hipGetLastError());
    if (loss) return read_loss(n, nblk, s, loss, (long long)B * n->nout);
    return 0;
}

int qcml_gradients(qcml_net* n, const float* dX, const float* dT, int32_t B, float* host_grad, float* loss)
{
    if (!n || !dX || !dT || !host_grad || B < 1 || B > n->maxb) return 1;
    HIPCHK(hipSetDevice(n->device));
    int nblk = 0;
    int rc = backward(n, dX, dT, B, (hipStream_t)0, &nblk);
    if (rc) return rc;
    if (loss) {
        rc = read_loss(n, nblk, (hipStream_t)0, loss, (long long)B * n->nout);
        if (rc) return rc;
    }
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(host_grad, n->g, n->nparam * sizeof(float), hipMemcpyDeviceToHost));
    return 0;
}

int qcml_gemm(int32_t M, int32_t N, int32_t K, const void* A, int32_t a_bf16, int64_t sam, int64_t sak, const void* B,
              int32_t b_bf16, int64_t sbk, int64_t sbn, float* C, int64_t ldc, float alpha, float beta, void* stream)
{
    if (!A || !B || !C || M < 0 || N < 0 || K < 0) return 1;
    HIPCHK(gemm(M, N, K, A, a_bf16 != 0, sam, sak, B, b_bf16 != 0, sbk, sbn, C, ldc, alpha, beta, (hipStream_t)stream));
    return 0;
}

void qcml_destroy(qcml_net* n) { free_net(n); }

}  // extern "C"
