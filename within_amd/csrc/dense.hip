// dense.hip -- the anchored dense-core inverse of the approximate Cholesky
// factor on the GPU (operator/local_solver.rs:207-300 semantics), used at
// setup for large cores.
//
// The core's Laplacian arrives as CSR (the host assembles no dense matrix);
// the arithmetic of every entry is the one fixed by anchored_inverse in
// plan.cpp (and by the oracle's plain loops):
//   l_ij = (a_ij - sum_{k<j} l_ik l_jk) / l_jj,  l_jj = sqrt(a_jj - sum_{k<j} l_jk^2)
//   U_ij = (delta_ij - sum_{k=j}^{i-1} l_ik U_kj) / l_ii         (U = L^-1)
//   inv_ij = sum_{k>=i} U_ki U_kj                                  (j <= i)
// each sum term by term, k ascending, one rounded multiply and one rounded
// add or subtract per term (this file is built without FP contraction).  The
// kernels only choose which thread evaluates which entry and where operands
// are staged, so the result is the host factorisation bit for bit.  Terms
// whose U factor is an exact zero (k below the diagonal of U) are added where
// that keeps the loops uniform; they leave a +0 or nonzero partial sum
// unchanged.
//
// Stage 1 is left-looking by 32-column tiles: the tile's columns receive the
// contributions of the earlier columns, then the diagonal tile is finished
// column by column and the rows below it are substituted against it (one
// thread per row).  Stage 2 runs row tiles in order, each column tile of the
// row tile in its own workgroup, then the tile's own rows in order.  In both,
// every 1024 columns (rows) a bulk launch adds the terms of the columns (rows)
// finished since the previous one to the whole trailing part, so a tile's own
// update only sums the terms since the last bulk launch: the late tiles, with
// few rows left, no longer run long serial sums in few workgroups.  Stage 3
// evaluates the whole lower triangle at once.  Every entry still receives its
// terms in ascending k; partial sums written to memory and read back are
// exact, so splitting a sum across kernels keeps its rounding.
//
// Everything runs on a stream of its own.  Once the Cholesky is known to have
// succeeded, the triangular inverse and the product are left running: the
// returned inverse holds the stream and the scratch, and DeviceDense::ready()
// waits for them before the matrix is read.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "plan.hpp"

namespace wa {
namespace {

constexpr int TB = 32;  // column tile
constexpr int RB = 64;  // rows per workgroup (one thread each)
constexpr int KC = 32;  // k values staged in LDS per step

struct HipFail : Error {
  explicit HipFail(const std::string& m) : Error(10, m) {}
};
#define DCHK(x)                                                                 \
  do {                                                                          \
    hipError_t e_ = (x);                                                        \
    if (e_ != hipSuccess) throw HipFail(std::string("HIP: ") + hipGetErrorString(e_)); \
  } while (0)

// Left-looking update of the column tile [c0, c0 + w) for rows >= c0:
// L_ij -= sum_{ka<=k<c0} L_ik L_jk, k ascending, continuing the stored partial
// sum (k < ka came from the bulk updates; the tile kernels continue from
// k = c0).  A workgroup takes 32 rows, each thread a 2 x 2 block of entries
// whose sums it runs term by term; the operands of a 32-wide k chunk are
// staged in LDS k-major, and the next
// chunk's are loaded while one is summed.
constexpr int UR = 32;  // rows per update workgroup
constexpr int kSwitch = 1024;  // columns between bulk updates (a multiple of 64)
__global__ void __launch_bounds__(256) k_chol_update(double* __restrict__ L, int m, int c0, int w, int ka) {
  __shared__ double As[KC][UR + 2];
  __shared__ double Bs[KC][TB + 2];
  const int t = threadIdx.x, tr = t / 16, tc = t % 16;
  const int64_t ib = c0 + (int64_t)blockIdx.x * UR;
  double acc[2][2];
  bool live[2][2];
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y) {
      const int64_t i = ib + 2 * tr + x;
      const int jl = 2 * tc + y;
      live[x][y] = i < m && jl < w && c0 + jl <= i;
      acc[x][y] = live[x][y] ? L[i * m + c0 + jl] : 0.0;
    }
  constexpr int NS = UR * KC / 256;
  double an[NS], bn[NS];
  auto fetch = [&](int kc) {
    const int kw = min(KC, c0 - kc);
#pragma unroll
    for (int u = 0; u < NS; ++u) {
      const int e = t + 256 * u, r = e / KC, kk = e % KC;
      an[u] = (ib + r < m && kk < kw) ? L[(ib + r) * m + kc + kk] : 0.0;
      bn[u] = (r < w && kk < kw) ? L[(int64_t)(c0 + r) * m + kc + kk] : 0.0;
    }
  };
  if (ka < c0) fetch(ka);
  for (int kc = ka; kc < c0; kc += KC) {
    const int kw = min(KC, c0 - kc);
#pragma unroll
    for (int u = 0; u < NS; ++u) {
      const int e = t + 256 * u;
      As[e % KC][e / KC] = an[u];
      Bs[e % KC][e / KC] = bn[u];
    }
    __syncthreads();
    if (kc + KC < c0) fetch(kc + KC);
    for (int kk = 0; kk < kw; ++kk) {
      const double a0 = As[kk][2 * tr], a1 = As[kk][2 * tr + 1];
      const double b0 = Bs[kk][2 * tc], b1 = Bs[kk][2 * tc + 1];
      acc[0][0] = acc[0][0] - a0 * b0;
      acc[0][1] = acc[0][1] - a0 * b1;
      acc[1][0] = acc[1][0] - a1 * b0;
      acc[1][1] = acc[1][1] - a1 * b1;
    }
    __syncthreads();
  }
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y)
      if (live[x][y]) L[(ib + 2 * tr + x) * m + c0 + 2 * tc + y] = acc[x][y];
}

// Diagonal tile [c0, c0 + w) after its update: rows c0 + t, t < w, finished
// column by column (k from c0 on).
__global__ void __launch_bounds__(64) k_chol_diag(double* __restrict__ L, int m, int c0, int w,
                                                  int* __restrict__ fail) {
  __shared__ double R[TB][TB + 1];
  const int t = threadIdx.x;
  for (int e = t; e < TB * TB; e += 64) {
    const int r = e / TB, k = e % TB;
    R[r][k] = (r < w && k <= r) ? L[(int64_t)(c0 + r) * m + c0 + k] : 0.0;
  }
  __syncthreads();
  for (int s = 0; s < w; ++s) {
    if (t == s) {
      double d = R[s][s];
      for (int k = 0; k < s; ++k) d = d - R[s][k] * R[s][k];
      if (!(d > 0.0)) *fail = 1;
      R[s][s] = __dsqrt_rn(d);
    }
    __syncthreads();
    if (t > s && t < w) {
      double x = R[t][s];
      for (int k = 0; k < s; ++k) x = x - R[t][k] * R[s][k];
      R[t][s] = x / R[s][s];
    }
    __syncthreads();
  }
  for (int e = t; e < TB * TB; e += 64) {
    const int r = e / TB, k = e % TB;
    if (r < w && k <= r) L[(int64_t)(c0 + r) * m + c0 + k] = R[r][k];
  }
}

// Rows r0 + blockIdx.x * RB + t (< m) of the column tile [c0, c0 + w) after
// its update: forward substitution against the finished diagonal tile.
__global__ void __launch_bounds__(RB) k_chol_below(double* __restrict__ L, int m, int c0, int w, int r0) {
  __shared__ double As[RB][TB + 1];
  __shared__ double D[TB][TB + 1];
  const int t = threadIdx.x;
  const int64_t ib = r0 + (int64_t)blockIdx.x * RB, i = ib + t;
  for (int e = t; e < RB * TB; e += RB) {
    const int r = e / TB, k = e % TB;
    As[r][k] = (ib + r < m && k < w) ? L[(ib + r) * m + c0 + k] : 0.0;
  }
  for (int e = t; e < TB * TB; e += RB) {
    const int r = e / TB, k = e % TB;
    D[r][k] = (r < w && k <= r) ? L[(int64_t)(c0 + r) * m + c0 + k] : 1.0;
  }
  __syncthreads();
  double* r_ = As[t];
  for (int jj = 0; jj < w; ++jj) {
    double x = r_[jj];
    for (int k = 0; k < jj; ++k) x = x - r_[k] * D[jj][k];
    r_[jj] = x / D[jj][jj];
  }
  if (i < m)
    for (int jj = 0; jj < w; ++jj) L[i * m + c0 + jj] = r_[jj];
}

// U rows [r0, r0 + TB) of column tile j0 = blockIdx.x * TB (<= r0), from the
// partial sums stored in U (delta_ij less the terms k < ka of the bulk
// updates): each thread a 2 x 2 block of entries for k in [max(j0, ka), r0)
// (term by term, operands staged k-major in LDS, the next chunk loaded while
// one is summed), then the tile's own rows in order, one thread per column.
__global__ void __launch_bounds__(256) k_inv_lower(const double* __restrict__ L, double* __restrict__ U, int m,
                                                   int r0, int ka) {
  __shared__ double Ls[KC][TB + 2];
  __shared__ double Us[KC][TB + 2];
  __shared__ double S[TB][TB + 1];
  __shared__ double Lt[TB][TB + 1];
  const int t = threadIdx.x, tr = t / 16, tc = t % 16;
  const int j0 = blockIdx.x * TB;
  double acc[2][2];
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y) {
      const int i = r0 + 2 * tr + x, j = j0 + 2 * tc + y;
      acc[x][y] = (i < m && j < m) ? U[(int64_t)i * m + j] : 0.0;
    }
  constexpr int NS = TB * KC / 256;
  double ln[NS], un[NS];
  auto fetch = [&](int kc) {
    const int kw = min(KC, r0 - kc);
#pragma unroll
    for (int u = 0; u < NS; ++u) {
      const int e = t + 256 * u;
      const int lr = e / KC, lk = e % KC, k2 = e / TB, jj = e % TB;
      ln[u] = (r0 + lr < m && lk < kw) ? L[(int64_t)(r0 + lr) * m + kc + lk] : 0.0;
      un[u] = (k2 < kw && j0 + jj < m) ? U[(int64_t)(kc + k2) * m + j0 + jj] : 0.0;
    }
  };
  const int k0 = max(j0, ka);
  if (k0 < r0) fetch(k0);
  for (int kc = k0; kc < r0; kc += KC) {
    const int kw = min(KC, r0 - kc);
#pragma unroll
    for (int u = 0; u < NS; ++u) {
      const int e = t + 256 * u;
      Ls[e % KC][e / KC] = ln[u];
      Us[e / TB][e % TB] = un[u];
    }
    __syncthreads();
    if (kc + KC < r0) fetch(kc + KC);
    for (int kk = 0; kk < kw; ++kk) {
      const double a0 = Ls[kk][2 * tr], a1 = Ls[kk][2 * tr + 1];
      const double b0 = Us[kk][2 * tc], b1 = Us[kk][2 * tc + 1];
      acc[0][0] = acc[0][0] - a0 * b0;
      acc[0][1] = acc[0][1] - a0 * b1;
      acc[1][0] = acc[1][0] - a1 * b0;
      acc[1][1] = acc[1][1] - a1 * b1;
    }
    __syncthreads();
  }
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y) S[2 * tr + x][2 * tc + y] = acc[x][y];
  for (int e = t; e < TB * TB; e += 256) {
    const int r = e / TB, c = e % TB;
    Lt[r][c] = (r0 + r < m && c <= r) ? L[(int64_t)(r0 + r) * m + r0 + c] : 1.0;
  }
  __syncthreads();
  // k in [r0, i): one thread per column, the column kept in S
  if (t < TB) {
    const int j = j0 + t;
    for (int s = 0; s < TB; ++s) {
      double x = S[s][t];
      for (int k = 0; k < s; ++k) x = x - Lt[s][k] * S[k][t];
      S[s][t] = (r0 + s >= j) ? x / Lt[s][s] : 0.0;
    }
  }
  __syncthreads();
  for (int e = t; e < TB * TB; e += 256) {
    const int r = e / TB, c = e % TB;
    const int64_t i = r0 + r;
    const int j = j0 + c;
    if (i < m && j < m) U[i * m + j] = S[r][c];
  }
}

// Bulk sums over 64 x 64 tiles, each thread a 4 x 4 block of entries whose
// terms it adds one by one, k ascending (operands of a 16-wide k chunk staged
// in LDS, the next chunk loaded while one is summed; partial sums stored in
// the output are continued and stored back, which is exact):
//   kChol: L_ij -= sum_{ka<=k<kb} L_ik L_jk over tiles i, j >= (i_lo, j_lo), j <= i
//   kTri:  U_ij -= sum_{ka<=k<kb} L_ik U_kj over rows >= i_lo, columns >= j_lo
//          (terms with k < j are exact zeros of U)
//   kSym:  inv_ij = inv_ji = sum_{k>=i0} U_ki U_kj, j <= i, i0 the tile's first row
//          (terms with k < i are exact zeros of U)
enum { kChol = 0, kTri = 1, kSym = 2 };
constexpr int GT = 64, GK = 16;
template <int MODE>
__global__ void __launch_bounds__(256) k_tile_terms(const double* __restrict__ A, const double* __restrict__ B,
                                                    double* __restrict__ C, int m, int ldc, int i_lo, int j_lo,
                                                    int ka, int kb) {
  __shared__ __align__(16) double As[GK][GT + 2];
  __shared__ __align__(16) double Bs[GK][GT + 2];
  const int t = threadIdx.x, tr = t / 16, tc = t % 16;
  const int i0 = i_lo + blockIdx.x * GT, j0 = j_lo + blockIdx.y * GT;
  if (MODE != kTri && j0 > i0 + GT - 1) return;
  if (MODE == kSym) ka = i0, kb = m;
  double acc[4][4];
  bool live[4][4];
#pragma unroll
  for (int x = 0; x < 4; ++x)
#pragma unroll
    for (int y = 0; y < 4; ++y) {
      const int i = i0 + 4 * tr + x, j = j0 + 4 * tc + y;
      live[x][y] = i < m && j < m && (MODE == kTri || j <= i);
      acc[x][y] = (MODE != kSym && live[x][y]) ? C[(int64_t)i * ldc + j] : 0.0;
    }
  constexpr int NS = GT * GK / 256;
  double an[NS], bn[NS];
  auto fetch = [&](int kc) {
    const int kw = min(GK, kb - kc);
#pragma unroll
    for (int u = 0; u < NS; ++u) {
      const int e = t + 256 * u;
      if (MODE == kSym) {
        const int kk = e / GT, r = e % GT;
        an[u] = (kk < kw && i0 + r < m) ? A[(int64_t)(kc + kk) * m + i0 + r] : 0.0;
      } else {
        const int r = e / GK, kk = e % GK;
        an[u] = (kk < kw && i0 + r < m) ? A[(int64_t)(i0 + r) * m + kc + kk] : 0.0;
      }
      if (MODE == kChol) {
        const int r = e / GK, kk = e % GK;
        bn[u] = (kk < kw && j0 + r < m) ? B[(int64_t)(j0 + r) * m + kc + kk] : 0.0;
      } else {
        const int kk = e / GT, c = e % GT;
        bn[u] = (kk < kw && j0 + c < m) ? B[(int64_t)(kc + kk) * m + j0 + c] : 0.0;
      }
    }
  };
  if (ka < kb) fetch(ka);
  for (int kc = ka; kc < kb; kc += GK) {
    const int kw = min(GK, kb - kc);
#pragma unroll
    for (int u = 0; u < NS; ++u) {
      const int e = t + 256 * u;
      if (MODE == kSym)
        As[e / GT][e % GT] = an[u];
      else
        As[e % GK][e / GK] = an[u];
      if (MODE == kChol)
        Bs[e % GK][e / GK] = bn[u];
      else
        Bs[e / GT][e % GT] = bn[u];
    }
    __syncthreads();
    if (kc + GK < kb) fetch(kc + GK);
    for (int kk = 0; kk < kw; ++kk) {
      double a[4], b[4];
#pragma unroll
      for (int x = 0; x < 4; ++x) a[x] = As[kk][4 * tr + x];
#pragma unroll
      for (int y = 0; y < 4; ++y) b[y] = Bs[kk][4 * tc + y];
#pragma unroll
      for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y) {
          if (MODE == kSym)
            acc[x][y] = acc[x][y] + a[x] * b[y];
          else
            acc[x][y] = acc[x][y] - a[x] * b[y];
        }
    }
    __syncthreads();
  }
#pragma unroll
  for (int x = 0; x < 4; ++x)
#pragma unroll
    for (int y = 0; y < 4; ++y)
      if (live[x][y]) {
        const int64_t i = i0 + 4 * tr + x, j = j0 + 4 * tc + y;
        C[i * ldc + j] = acc[x][y];
        if (MODE == kSym) C[j * ldc + i] = acc[x][y];
      }
}

// U = I on the diagonal (U is zeroed first): the starting partial sums delta_ij
__global__ void k_unit_diag(double* __restrict__ U, int64_t m) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x)
    U[i * m + i] = 1.0;
}

// The anchored minor's lower triangle (rows and columns [0, m), m = n - 1) of
// the n x n Laplacian given as CSR with positive off-diagonal weights and its
// diagonal: L[i][j] = -w_ij for j < i, L[i][i] = diag_i; L is zeroed first.
// One workgroup per row.
__global__ void k_minor_from_csr(const uint32_t* __restrict__ ptr, const uint32_t* __restrict__ col,
                                 const double* __restrict__ w, const double* __restrict__ diag,
                                 double* __restrict__ L, int64_t m) {
  const int64_t i = blockIdx.x;
  if (i >= m) return;
  for (uint32_t e = ptr[i] + threadIdx.x; e < ptr[i + 1]; e += blockDim.x) {
    const uint32_t j = col[e];
    if (j < i) L[i * m + j] = -w[e];
  }
  if (threadIdx.x == 0) L[i * m + i] = diag[i];
}

struct DevBuf {
  void* p = nullptr;
  explicit DevBuf(size_t bytes) { DCHK(hipMalloc(&p, bytes)); }
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  void* release() {
    void* r = p;
    p = nullptr;
    return r;
  }
  template <typename T>
  T* as() const { return (T*)p; }
};

}  // namespace

DeviceDense::~DeviceDense() {
  try {
    ready();
  } catch (const std::exception&) {
  }
  if (p) (void)hipFree(p);
}

void DeviceDense::ready() const {
  std::lock_guard<std::mutex> g(mu);
  if (!pending) return;
  hipStream_t s = static_cast<hipStream_t>(pending);
  pending = nullptr;
  const hipError_t e = hipStreamSynchronize(s);
  (void)hipStreamDestroy(s);
  for (void* b : scratch) (void)hipFree(b);
  scratch.clear();
  if (e != hipSuccess) throw HipFail(std::string("HIP: ") + hipGetErrorString(e));
}

std::vector<double> DeviceDense::download() const {
  ready();
  std::vector<double> h((size_t)n * n);
  if (!h.empty()) DCHK(hipMemcpy(h.data(), p, sizeof(double) * h.size(), hipMemcpyDeviceToHost));
  return h;
}

bool device_anchored_inverse(i64 n, const u32* ptr, const u32* col, const double* w, const double* diag,
                             std::shared_ptr<DeviceDense>& out) {
  const bool report = std::getenv("WITHIN_AMD_VERBOSE") != nullptr;
  auto t0 = std::chrono::steady_clock::now();
  auto lap = [&](const char* what) {
    if (!report) return;
    (void)hipDeviceSynchronize();
    const auto t1 = std::chrono::steady_clock::now();
    std::fprintf(stderr, "[within_amd]       dense core %s: %.3f s\n", what,
                 std::chrono::duration<double>(t1 - t0).count());
    t0 = t1;
  };
  int dev = 0;
  DCHK(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  DCHK(hipGetDeviceProperties(&prop, dev));
  if (std::string(prop.gcnArchName).find("gfx950") == std::string::npos)
    throw HipFail(std::string("within_amd requires a gfx950 (MI355X) device, found ") + prop.gcnArchName);
  // everything runs on a stream of its own (not ordered with the null stream,
  // so the caller's uploads proceed meanwhile); the stream is synchronised
  // and destroyed on every way out except the last, where the inverse takes it
  struct Stream {
    hipStream_t s = nullptr;
    ~Stream() {
      if (s) {
        (void)hipStreamSynchronize(s);
        (void)hipStreamDestroy(s);
      }
    }
  } ds;
  DCHK(hipStreamCreateWithFlags(&ds.s, hipStreamNonBlocking));
  const hipStream_t st = ds.s;
  const size_t nn = (size_t)n * n * sizeof(double);
  auto inv = std::make_shared<DeviceDense>();
  inv->n = n;
  DCHK(hipMalloc(&inv->p, std::max<size_t>(nn, sizeof(double))));
  DCHK(hipMemsetAsync(inv->p, 0, std::max<size_t>(nn, sizeof(double)), st));
  const i64 m = n - 1;
  if (m <= 0) {
    DCHK(hipStreamSynchronize(st));
    out = inv;
    return true;
  }
  const size_t mm = (size_t)m * m * sizeof(double);
  DevBuf dL(mm), dU(mm), dF(sizeof(int));
  // on a way out with work in flight, it finishes before any buffer is freed
  struct SyncOnExit {
    hipStream_t* s;
    ~SyncOnExit() {
      if (*s) (void)hipStreamSynchronize(*s);
    }
  } sync_exit{&ds.s};
  {
    // the sparse Laplacian goes up; the minor's dense lower triangle is laid out on the device
    const size_t nnz = ptr[n];
    DevBuf dP(sizeof(u32) * (n + 1)), dC(sizeof(u32) * std::max<size_t>(nnz, 1)),
        dW(sizeof(double) * std::max<size_t>(nnz, 1)), dD(sizeof(double) * n);
    // on the core's stream, so the layout kernel is ordered after them (the
    // host arrays outlive the synchronisation below)
    DCHK(hipMemcpyAsync(dP.p, ptr, sizeof(u32) * (n + 1), hipMemcpyHostToDevice, st));
    if (nnz) {
      DCHK(hipMemcpyAsync(dC.p, col, sizeof(u32) * nnz, hipMemcpyHostToDevice, st));
      DCHK(hipMemcpyAsync(dW.p, w, sizeof(double) * nnz, hipMemcpyHostToDevice, st));
    }
    DCHK(hipMemcpyAsync(dD.p, diag, sizeof(double) * n, hipMemcpyHostToDevice, st));
    DCHK(hipMemsetAsync(dL.p, 0, mm, st));
    k_minor_from_csr<<<(unsigned)m, 256, 0, st>>>(dP.as<u32>(), dC.as<u32>(), dW.as<double>(), dD.as<double>(),
                                                   dL.as<double>(), m);
    DCHK(hipGetLastError());
    DCHK(hipStreamSynchronize(st));  // before the CSR buffers are freed
  }
  DCHK(hipMemsetAsync(dU.p, 0, mm, st));
  k_unit_diag<<<(unsigned)std::min<i64>((m + 255) / 256, 1024), 256, 0, st>>>(dU.as<double>(), m);
  DCHK(hipMemsetAsync(dF.p, 0, sizeof(int), st));
  lap("upload");
  const int mi = (int)m, ni = (int)n;
  // every kSwitch columns the terms of the finished columns since the last
  // switch are summed into the whole trailing part at once; each tile's own
  // update then sums only the columns since the last switch
  auto tiles = [](int len) { return (unsigned)((len + GT - 1) / GT); };
  int ka = 0;
  for (int c0 = 0; c0 < mi; c0 += TB) {
    const int w = std::min(TB, mi - c0), c1 = c0 + w;
    if (c0 > 0 && c0 % kSwitch == 0) {
      k_tile_terms<kChol><<<dim3(tiles(mi - c0), tiles(mi - c0)), 256, 0, st>>>(
          dL.as<double>(), dL.as<double>(), dL.as<double>(), mi, mi, c0, c0, ka, c0);
      ka = c0;
    }
    if (c0 > ka) k_chol_update<<<(unsigned)((mi - c0 + UR - 1) / UR), 256, 0, st>>>(dL.as<double>(), mi, c0, w, ka);
    k_chol_diag<<<1, 64, 0, st>>>(dL.as<double>(), mi, c0, w, dF.as<int>());
    if (c1 < mi) k_chol_below<<<(unsigned)((mi - c1 + RB - 1) / RB), RB, 0, st>>>(dL.as<double>(), mi, c0, w, c1);
  }
  DCHK(hipGetLastError());
  int failed = 0;
  DCHK(hipMemcpyAsync(&failed, dF.p, sizeof(int), hipMemcpyDeviceToHost, st));
  DCHK(hipStreamSynchronize(st));
  if (failed) return false;
  lap("cholesky");
  ka = 0;
  for (int r0 = 0; r0 < mi; r0 += TB) {
    if (r0 > 0 && r0 % kSwitch == 0) {
      k_tile_terms<kTri><<<dim3(tiles(mi - r0), tiles(r0)), 256, 0, st>>>(dL.as<double>(), dU.as<double>(),
                                                                          dU.as<double>(), mi, mi, r0, 0, ka, r0);
      ka = r0;
    }
    k_inv_lower<<<(unsigned)(r0 / TB + 1), 256, 0, st>>>(dL.as<double>(), dU.as<double>(), mi, r0, ka);
  }
  DCHK(hipGetLastError());
  lap("triangular inverse");
  k_tile_terms<kSym><<<dim3(tiles(mi), tiles(mi)), 256, 0, st>>>(dU.as<double>(), dU.as<double>(),
                                                                 static_cast<double*>(inv->p), mi, ni, 0, 0, 0, 0);
  DCHK(hipGetLastError());
  lap("product");
  // the triangular inverse and the product finish behind the caller's back:
  // the inverse takes the stream and the two scratch matrices (ready() waits)
  {
    std::lock_guard<std::mutex> g(inv->mu);
    inv->scratch = {dL.release(), dU.release()};
    inv->pending = ds.s;
    ds.s = nullptr;
  }
  out = inv;
  return true;
}

}  // namespace wa
