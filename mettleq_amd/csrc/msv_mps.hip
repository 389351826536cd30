// Matrix-product-state engine on one MI355X (gfx950): the reference's
// MPSState (mps_state.py:250-1549, native core _mps_native.cpp) with every
// site tensor resident in HBM. Tensor A[i] has shape (dl, 2, dr), stored
// row-major (index (a * 2 + s) * dr + b), complex64 like the reference's
// storage. Site i is qubit i (qubit 0 = the most significant amplitude bit).
//
// Per two-site update (sites i, i + 1):
//   k_mps_merge   theta = G . (A_i A_{i+1}), written column-major as the
//                 (2 dl) x (2 dr) matrix the SVD takes;
//   k_mps_svd     one-sided Jacobi SVD in one workgroup on the device (the
//                 reference: LAPACK cgesdd on the CPU);
//   host           the reference's truncation rule on the singular values
//                 (mps_state.py:193-247), read back (2 min(dl, dr) floats);
//   k_mps_split    A_i = U[:, :r], A_{i+1} = diag(S / norm) Vh[:r, :]
//                 (or U diag(S) and Vh when the center moves left).
// The orthogonality center moves like the reference's QR moves
// (mps_state.py:347-399): one-site SVDs without truncation, the factor
// carried into the neighbour by k_mps_absorb_*.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mettleq_sv.h"
#include "msv_internal.h"

using msv::set_error;

namespace {

using C = float2;

__device__ __forceinline__ C cmul(C a, C b) { return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ __forceinline__ C cadd(C a, C b) { return make_float2(a.x + b.x, a.y + b.y); }

struct Gate2 {
  C u[4];
};
struct Gate4 {
  C g[16];  // row-major, basis index 2 s_left + s_right
};

// A[a, s, b] <- sum_t U[s, t] A[a, t, b] for one site, in place.
__global__ void __launch_bounds__(256) k_mps_1q(C *A, int dl, int dr, Gate2 U) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= dl * dr) return;
  const int a = idx / dr, b = idx % dr;
  C *p0 = A + (a * 2) * dr + b;
  C *p1 = A + (a * 2 + 1) * dr + b;
  const C x0 = *p0, x1 = *p1;
  *p0 = cadd(cmul(U.u[0], x0), cmul(U.u[1], x1));
  *p1 = cadd(cmul(U.u[2], x0), cmul(U.u[3], x1));
}

// theta[(a, s), (t, b)] = sum_{s0, t0} G[2 s + t, 2 s0 + t0] sum_k L[a, s0, k] R[k, t0, b],
// column-major with m = 2 dl rows: row a * 2 + s, column t * dr + b.
__global__ void __launch_bounds__(256) k_mps_merge(const C *__restrict__ L, const C *__restrict__ R,
                                                   C *__restrict__ T, int dl, int dm, int dr, Gate4 G) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= dl * dr) return;
  const int a = idx / dr, b = idx % dr;
  C t[4] = {make_float2(0, 0), make_float2(0, 0), make_float2(0, 0), make_float2(0, 0)};
  for (int k = 0; k < dm; ++k) {
    const C l0 = L[(a * 2) * dm + k], l1 = L[(a * 2 + 1) * dm + k];
    const C r0 = R[(k * 2) * dr + b], r1 = R[(k * 2 + 1) * dr + b];
    t[0] = cadd(t[0], cmul(l0, r0));
    t[1] = cadd(t[1], cmul(l0, r1));
    t[2] = cadd(t[2], cmul(l1, r0));
    t[3] = cadd(t[3], cmul(l1, r1));
  }
  const int m = 2 * dl;
#pragma unroll
  for (int o = 0; o < 4; ++o) {
    C v = make_float2(0, 0);
#pragma unroll
    for (int i = 0; i < 4; ++i) v = cadd(v, cmul(G.g[o * 4 + i], t[i]));
    const int s = o >> 1, tt = o & 1;
    T[(a * 2 + s) + static_cast<size_t>(tt * dr + b) * m] = v;
  }
}

// Site tensor viewed as a (rows x cols) row-major matrix, copied into a
// column-major one (the SVD's input).
__global__ void __launch_bounds__(256) k_mps_to_colmajor(const C *__restrict__ A, C *__restrict__ M, int rows,
                                                         int cols) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= rows * cols) return;
  const int r = idx / cols, c = idx % cols;
  M[r + static_cast<size_t>(c) * rows] = A[idx];
}

// Split after a two-site SVD: U (m x kmin, column-major), S, Vh (kmin x n,
// column-major, ldv = kmin). left (dl, 2, r) = U[:, :r] (x S / norm when the
// singular values go left), right (r, 2, dr) = (S / norm) Vh[:r, :] (or Vh).
__global__ void __launch_bounds__(256) k_mps_split(const C *__restrict__ U, const float *__restrict__ S,
                                                   const C *__restrict__ Vh, C *__restrict__ left,
                                                   C *__restrict__ right, int dl, int dr, int kmin, int r,
                                                   float scale, int s_left) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  const int m = 2 * dl, n = 2 * dr;
  const int nl = m * r, nr = r * n;
  if (idx < nl) {
    const int row = idx / r, j = idx % r;  // row = a * 2 + s
    C u = U[row + static_cast<size_t>(j) * m];
    if (s_left) {
      const float f = S[j] * scale;
      u.x *= f;
      u.y *= f;
    }
    left[idx] = u;
  } else if (idx < nl + nr) {
    const int k = idx - nl;
    const int j = k / n, col = k % n;  // col = t * dr + b; right index (j * 2 + t) * dr + b = j * n + col
    C v = Vh[j + static_cast<size_t>(col) * kmin];
    if (!s_left) {
      const float f = S[j] * scale;
      v.x *= f;
      v.y *= f;
    }
    right[k] = v;
  }
}

// Center moves right: A_i <- U (rows x r, from the column-major U), and the
// next site absorbs X = diag(S) Vh (r x cols): next'[j, t, b] = sum_c X[j, c] next[c, t, b].
__global__ void __launch_bounds__(256) k_mps_absorb_right(const C *__restrict__ U, const float *__restrict__ S,
                                                          const C *__restrict__ Vh, C *__restrict__ site,
                                                          const C *__restrict__ next, C *__restrict__ next_out,
                                                          int rows, int cols, int kmin, int r, int dn) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  const int nu = rows * r, nn = r * 2 * dn;
  if (idx < nu) {
    const int row = idx / r, j = idx % r;
    site[idx] = U[row + static_cast<size_t>(j) * rows];
  } else if (idx < nu + nn) {
    const int k = idx - nu;
    const int j = k / (2 * dn), tb = k % (2 * dn);
    C acc = make_float2(0, 0);
    for (int c = 0; c < cols; ++c) acc = cadd(acc, cmul(Vh[j + static_cast<size_t>(c) * kmin], next[c * 2 * dn + tb]));
    acc.x *= S[j];
    acc.y *= S[j];
    next_out[k] = acc;
  }
}

// Center moves left: A_i (rows = dl) x (cols = 2 dr) = U S Vh; A_i <- Vh[:r, :]
// and the previous site absorbs Y = U diag(S) (rows x r):
// prev'[a, t, j] = sum_c prev[a, t, c] Y[c, j].
__global__ void __launch_bounds__(256) k_mps_absorb_left(const C *__restrict__ U, const float *__restrict__ S,
                                                         const C *__restrict__ Vh, C *__restrict__ site,
                                                         const C *__restrict__ prev, C *__restrict__ prev_out,
                                                         int rows, int cols, int kmin, int r, int dp) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  const int nv = r * cols, np = dp * 2 * r;
  if (idx < nv) {
    const int j = idx / cols, c = idx % cols;
    site[idx] = Vh[j + static_cast<size_t>(c) * kmin];
  } else if (idx < nv + np) {
    const int k = idx - nv;
    const int at = k / r, j = k % r;  // at = a * 2 + t
    C acc = make_float2(0, 0);
    for (int c = 0; c < rows; ++c) acc = cadd(acc, cmul(prev[at * rows + c], U[c + static_cast<size_t>(j) * rows]));
    acc.x *= S[j];
    acc.y *= S[j];
    prev_out[k] = acc;
  }
}

// One contraction step of the dense state: out[(p, s), b] = scale sum_a in[p, a] A[a, s, b].
__global__ void __launch_bounds__(256) k_mps_contract(const C *__restrict__ in, const C *__restrict__ A,
                                                      C *__restrict__ out, uint64_t np, int dm, int dn,
                                                      float scale) {
  const uint64_t idx = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const uint64_t total = np * 2 * static_cast<uint64_t>(dn);
  if (idx >= total) return;
  const uint64_t p = idx / (2 * dn);
  const int sb = static_cast<int>(idx % (2 * dn));
  C acc = make_float2(0, 0);
  for (int a = 0; a < dm; ++a) acc = cadd(acc, cmul(in[p * dm + a], A[a * 2 * dn + sb]));
  acc.x *= scale;
  acc.y *= scale;
  out[idx] = acc;
}

// Transfer step of <psi| O_1 x ... x O_n |psi>: env' = sum_{a, a', s, t}
// conj(A[a, s, b]) env[a, a'] O[s, t] A[a', t, b'] (complex128 environment).
__global__ void __launch_bounds__(256) k_mps_transfer(const double2 *__restrict__ env, const C *__restrict__ A,
                                                      double2 *__restrict__ out, int dm, int dn, Gate2 O) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= dn * dn) return;
  const int b = idx / dn, bp = idx % dn;
  double re = 0, im = 0;
  for (int a = 0; a < dm; ++a) {
    for (int s = 0; s < 2; ++s) {
      const C ca = A[(a * 2 + s) * dn + b];  // conj taken below
      // y = sum_{a', t} env[a, a'] O[s, t] A[a', t, b']
      double yr = 0, yi = 0;
      for (int ap = 0; ap < dm; ++ap) {
        const double2 e = env[a * dm + ap];
        const C x0 = A[(ap * 2) * dn + bp], x1 = A[(ap * 2 + 1) * dn + bp];
        const C o0 = O.u[s * 2], o1 = O.u[s * 2 + 1];
        const double tr = (double)o0.x * x0.x - (double)o0.y * x0.y + (double)o1.x * x1.x - (double)o1.y * x1.y;
        const double ti = (double)o0.x * x0.y + (double)o0.y * x0.x + (double)o1.x * x1.y + (double)o1.y * x1.x;
        yr += e.x * tr - e.y * ti;
        yi += e.x * ti + e.y * tr;
      }
      re += (double)ca.x * yr + (double)ca.y * yi;
      im += (double)ca.x * yi - (double)ca.y * yr;
    }
  }
  out[idx] = make_double2(re, im);
}

// One-sided Jacobi (Hestenes) SVD of the column-major rows x cols matrix A
// in one 1024-thread workgroup. With k = min(rows, cols) it works on W = A
// (or A^H when cols > rows, so that W has k columns of kk = max rows each)
// and V = I (k x k): each sweep visits every column pair (p, q) once, in
// k - 1 rounds of disjoint pairs (round-robin order, one wavefront per
// pair), and rotates the pair so that w_p^H w_q = 0 (phase e^{i phi} of
// w_p^H w_q, then the real Jacobi rotation t = sign(z) / (|z| + sqrt(1 + z^2)),
// z = (|w_q|^2 - |w_p|^2) / (2 |w_p^H w_q|)). Dot products accumulate in
// double precision. Sweeps stop when no pair is rotated (|w_p^H w_q| below
// kJacobiTol |w_p| |w_q|). Pairs with a column of norm below kJacobiFloor
// |A|_F are left alone: such a column's entries may be float denormals,
// whose few significant bits cannot reach kJacobiTol (the sweeps would
// rotate it forever), and its singular value lies below the exact-zero
// floor of the truncation, so its vector is never kept by a split. Then
// sigma_j = |w_j|, sorted descending:
//   A = U diag(S) Vh,  U (rows x k, column-major), Vh (k x cols, column-major,
//   leading dimension k); info = sweeps run, or -1 without convergence.
constexpr double kJacobiTol = 3e-7;
constexpr double kJacobiFloor = 1e-32;
constexpr int kJacobiMaxSweeps = 60;
constexpr int kSvdThreads = 1024;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__global__ void __launch_bounds__(kSvdThreads) k_mps_svd(const C *__restrict__ A, int rows, int cols,
                                                         C *__restrict__ W, C *__restrict__ V, C *__restrict__ U,
                                                         float *__restrict__ S, C *__restrict__ Vh,
                                                         int *__restrict__ info) {
  __shared__ int rotated;
  __shared__ double wave_fro[kSvdThreads / 64];
  __shared__ float sig[2048];
  __shared__ int rank[2048];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nwave = blockDim.x >> 6;
  const bool trans = cols > rows;
  const int kk = trans ? cols : rows;  // rows of W
  const int k = trans ? rows : cols;   // columns of W (= min(rows, cols))
  double fro = 0;
  for (int idx = tid; idx < kk * k; idx += blockDim.x) {
    const int i = idx % kk, j = idx / kk;
    C a;
    if (!trans) {
      a = A[i + static_cast<size_t>(j) * rows];
    } else {
      a = A[j + static_cast<size_t>(i) * rows];  // W = A^H: W[i, j] = conj(A[j, i])
      a.y = -a.y;
    }
    W[idx] = a;
    fro += (double)a.x * a.x + (double)a.y * a.y;
  }
  fro = wave_sum(fro);
  if (lane == 0) wave_fro[wave] = fro;
  for (int idx = tid; idx < k * k; idx += blockDim.x)
    V[idx] = make_float2((idx % k) == (idx / k) ? 1.f : 0.f, 0.f);
  __syncthreads();
  fro = 0;
  for (int w = 0; w < nwave; ++w) fro += wave_fro[w];
  const double floor2 = fmax(kJacobiFloor * kJacobiFloor * fro, 1e-66);
  const int np = k + (k & 1);  // players of the round-robin (one bye when k is odd)
  int sweeps = 0;
  bool converged = k < 2;
  for (; sweeps < kJacobiMaxSweeps && !converged; ++sweeps) {
    if (tid == 0) rotated = 0;
    __syncthreads();
    for (int r = 0; r < np - 1; ++r) {
      for (int pi = wave; pi < np / 2; pi += nwave) {
        // circle method: position 0 fixed, the others rotate by r
        const int a = pi == 0 ? 0 : 1 + (pi - 1 + r) % (np - 1);
        const int b = 1 + (np - 2 - pi + r) % (np - 1);
        const int p = a < b ? a : b, q = a < b ? b : a;
        if (q >= k) continue;
        C *wp = W + static_cast<size_t>(p) * kk, *wq = W + static_cast<size_t>(q) * kk;
        double al = 0, be = 0, gr = 0, gi = 0;
        for (int i = lane; i < kk; i += 64) {
          const C x = wp[i], y = wq[i];
          al += (double)x.x * x.x + (double)x.y * x.y;
          be += (double)y.x * y.x + (double)y.y * y.y;
          gr += (double)x.x * y.x + (double)x.y * y.y;  // conj(x) y
          gi += (double)x.x * y.y - (double)x.y * y.x;
        }
        al = wave_sum(al);
        be = wave_sum(be);
        gr = wave_sum(gr);
        gi = wave_sum(gi);
        const double g = sqrt(gr * gr + gi * gi);
        if (!(g > kJacobiTol * sqrt(al * be)) || !(fmin(al, be) > floor2)) continue;
        const double z = (be - al) / (2.0 * g);
        const double t = (z >= 0 ? 1.0 : -1.0) / (fabs(z) + sqrt(1.0 + z * z));
        const double c = 1.0 / sqrt(1.0 + t * t), sn = c * t;
        const double er = gr / g, ei = -gi / g;  // e^{-i phi}
        // w_p' = c w_p - s e^{-i phi} w_q,  w_q' = s w_p + c e^{-i phi} w_q
        const float fc = (float)c, fs = (float)sn;
        const float fr = (float)er, fi = (float)ei;
        for (int i = lane; i < kk; i += 64) {
          const C x = wp[i], y = wq[i];
          const C ey = make_float2(fr * y.x - fi * y.y, fr * y.y + fi * y.x);
          wp[i] = make_float2(fc * x.x - fs * ey.x, fc * x.y - fs * ey.y);
          wq[i] = make_float2(fs * x.x + fc * ey.x, fs * x.y + fc * ey.y);
        }
        C *vp = V + static_cast<size_t>(p) * k, *vq = V + static_cast<size_t>(q) * k;
        for (int i = lane; i < k; i += 64) {
          const C x = vp[i], y = vq[i];
          const C ey = make_float2(fr * y.x - fi * y.y, fr * y.y + fi * y.x);
          vp[i] = make_float2(fc * x.x - fs * ey.x, fc * x.y - fs * ey.y);
          vq[i] = make_float2(fs * x.x + fc * ey.x, fs * x.y + fc * ey.y);
        }
        if (lane == 0) rotated = 1;
      }
      __syncthreads();
    }
    converged = rotated == 0;
    __syncthreads();
  }
  // singular values (column norms) and their descending order
  for (int j = wave; j < k; j += nwave) {
    double s2 = 0;
    for (int i = lane; i < kk; i += 64) {
      const C x = W[static_cast<size_t>(j) * kk + i];
      s2 += (double)x.x * x.x + (double)x.y * x.y;
    }
    s2 = wave_sum(s2);
    if (lane == 0) sig[j] = (float)sqrt(s2);
  }
  __syncthreads();
  for (int j = tid; j < k; j += blockDim.x) {
    int rk = 0;
    for (int l = 0; l < k; ++l) rk += (sig[l] > sig[j] || (sig[l] == sig[j] && l < j)) ? 1 : 0;
    rank[j] = rk;
    S[rk] = sig[j];
  }
  __syncthreads();
  // outputs: the left vectors of W are w_j / sigma_j, its right vectors V
  for (int idx = tid; idx < kk * k; idx += blockDim.x) {
    const int i = idx % kk, j = idx / kk;
    const float inv = sig[j] > 0.f ? 1.f / sig[j] : 0.f;
    const C w = W[idx];
    const C u = make_float2(w.x * inv, w.y * inv);
    if (!trans) {
      U[i + static_cast<size_t>(rank[j]) * rows] = u;        // U[:, rank j] = w_j / sigma_j
    } else {
      Vh[rank[j] + static_cast<size_t>(i) * k] = make_float2(u.x, -u.y);  // Vh[rank j, :] = (w_j / sigma_j)^H
    }
  }
  for (int idx = tid; idx < k * k; idx += blockDim.x) {
    const int i = idx % k, j = idx / k;
    const C v = V[idx];
    if (!trans) {
      Vh[rank[j] + static_cast<size_t>(i) * k] = make_float2(v.x, -v.y);  // Vh[rank j, :] = v_j^H
    } else {
      U[i + static_cast<size_t>(rank[j]) * rows] = v;       // U[:, rank j] = v_j
    }
  }
  if (tid == 0) *info = converged ? sweeps : -1;
}

constexpr int kThreads = 256;
uint32_t blocks_for(uint64_t n) { return static_cast<uint32_t>((n + kThreads - 1) / kThreads); }

}  // namespace

struct msv_mps {
  int n = 0;
  int dmax = 64;
  double eps = 1e-10;
  bool renormalize = true;
  int device = 0;
  hipStream_t stream = nullptr;
  size_t cap = 0;                 // complex entries per site buffer (dmax x 2 x dmax)
  C *sites = nullptr;             // n site buffers
  C *spare = nullptr;             // one site buffer (out-of-place absorb)
  C *work = nullptr;              // SVD input / U / Vh / Jacobi W / Jacobi V (5 x 4 dmax^2)
  float *sv = nullptr;            // singular values and superdiagonal (2 x 2 dmax)
  int *info = nullptr;
  std::vector<int> dim;           // bond dimensions, dim[0] = dim[n] = 1
  int center = 0;
  msv_mps_stats st{};
  std::vector<float> h_sv;
};

namespace {

int fail_hip(hipError_t e, const char *what) {
  return set_error(MSV_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}
#define MPS_HIP(expr)                                 \
  do {                                                \
    const hipError_t e_ = (expr);                     \
    if (e_ != hipSuccess) return fail_hip(e_, #expr); \
  } while (0)
#define MPS_TRY(expr)              \
  do {                             \
    const int rc_ = (expr);        \
    if (rc_ != MSV_OK) return rc_; \
  } while (0)

int check_mps(const msv_mps *m) { return m ? MSV_OK : set_error(MSV_E_INVALID, "null MPS"); }
int check_site(const msv_mps *m, int q) {
  if (q < 0 || q >= m->n) return set_error(MSV_E_INVALID, "MPS qubit index out of range");
  return MSV_OK;
}
C *site_ptr(msv_mps *m, int i) { return m->sites + static_cast<size_t>(i) * m->cap; }

// The SVD of the (rows x cols) column-major matrix in work[0]: U at work[1]
// (rows x kmin), Vh at work[2] (kmin x cols), S on the device and in h_sv.
int svd(msv_mps *m, int rows, int cols) {
  const size_t slot = static_cast<size_t>(4) * m->dmax * m->dmax;
  const int kmin = std::min(rows, cols);
  hipLaunchKernelGGL(k_mps_svd, dim3(1), dim3(kSvdThreads), 0, m->stream, m->work, rows, cols, m->work + 3 * slot,
                     m->work + 4 * slot, m->work + slot, m->sv, m->work + 2 * slot, m->info);
  MPS_HIP(hipGetLastError());
  int h_info = 0;
  m->h_sv.resize(kmin);
  MPS_HIP(hipMemcpyAsync(m->h_sv.data(), m->sv, sizeof(float) * kmin, hipMemcpyDeviceToHost, m->stream));
  MPS_HIP(hipMemcpyAsync(&h_info, m->info, sizeof(int), hipMemcpyDeviceToHost, m->stream));
  MPS_HIP(hipStreamSynchronize(m->stream));
  m->st.svd_calls += 1;
  if (h_info < 0) return set_error(MSV_E_INVALID, "MPS SVD did not converge");
  for (int k = 0; k < kmin; ++k)
    if (!std::isfinite(m->h_sv[k])) return set_error(MSV_E_INVALID, "MPS SVD produced non-finite singular values");
  return MSV_OK;
}

// Moves the orthogonality center one site right (from i to i + 1) or left
// (from i to i - 1) without truncation: the reference's QR moves
// (mps_state.py:347-389), here through a one-site SVD.
int move_right(msv_mps *m, int i) {
  const int dl = m->dim[i], dr = m->dim[i + 1], dn = m->dim[i + 2];
  const int rows = 2 * dl, cols = dr, kmin = std::min(rows, cols);
  MPS_TRY(m->n > 0 ? MSV_OK : MSV_E_INVALID);
  hipLaunchKernelGGL(k_mps_to_colmajor, dim3(blocks_for(static_cast<uint64_t>(rows) * cols)), dim3(kThreads), 0,
                     m->stream, site_ptr(m, i), m->work, rows, cols);
  MPS_HIP(hipGetLastError());
  MPS_TRY(svd(m, rows, cols));
  const size_t slot = static_cast<size_t>(4) * m->dmax * m->dmax;
  const int r = kmin;
  hipLaunchKernelGGL(k_mps_absorb_right, dim3(blocks_for(static_cast<uint64_t>(rows) * r + r * 2 * dn)),
                     dim3(kThreads), 0, m->stream, m->work + slot, m->sv, m->work + 2 * slot, site_ptr(m, i),
                     site_ptr(m, i + 1), m->spare, rows, cols, kmin, r, dn);
  MPS_HIP(hipGetLastError());
  MPS_HIP(hipMemcpyAsync(site_ptr(m, i + 1), m->spare, sizeof(C) * r * 2 * dn, hipMemcpyDeviceToDevice, m->stream));
  m->dim[i + 1] = r;
  m->center = i + 1;
  return MSV_OK;
}

int move_left(msv_mps *m, int i) {
  const int dl = m->dim[i], dr = m->dim[i + 1], dp = m->dim[i - 1];
  const int rows = dl, cols = 2 * dr, kmin = std::min(rows, cols);
  hipLaunchKernelGGL(k_mps_to_colmajor, dim3(blocks_for(static_cast<uint64_t>(rows) * cols)), dim3(kThreads), 0,
                     m->stream, site_ptr(m, i), m->work, rows, cols);
  MPS_HIP(hipGetLastError());
  MPS_TRY(svd(m, rows, cols));
  const size_t slot = static_cast<size_t>(4) * m->dmax * m->dmax;
  const int r = kmin;
  hipLaunchKernelGGL(k_mps_absorb_left, dim3(blocks_for(static_cast<uint64_t>(r) * cols + dp * 2 * r)),
                     dim3(kThreads), 0, m->stream, m->work + slot, m->sv, m->work + 2 * slot, site_ptr(m, i),
                     site_ptr(m, i - 1), m->spare, rows, cols, kmin, r, dp);
  MPS_HIP(hipGetLastError());
  MPS_HIP(hipMemcpyAsync(site_ptr(m, i - 1), m->spare, sizeof(C) * dp * 2 * r, hipMemcpyDeviceToDevice, m->stream));
  m->dim[i] = r;
  m->center = i - 1;
  return MSV_OK;
}

int move_center(msv_mps *m, int target) {
  while (m->center < target) MPS_TRY(move_right(m, m->center));
  while (m->center > target) MPS_TRY(move_left(m, m->center));
  return MSV_OK;
}

// Two-site update of sites (i, i + 1) with the 4x4 gate g (row-major,
// complex double): merge, SVD, the reference's truncation
// (mps_state.py:193-247: keep r = min(rank, max(1, min(dmax, #{s >= eps s_0}))),
// discarded weight = sum of the dropped s^2, singular values divided by the
// kept norm when renormalizing), split with the center at i + 1.
constexpr double kZeroFloor = 1e-30;  // relative singular value treated as an exact zero

int two_site(msv_mps *m, int i, const std::complex<double> *g, double eps) {
  MPS_TRY(move_center(m, i));
  const int dl = m->dim[i], dm = m->dim[i + 1], dr = m->dim[i + 2];
  Gate4 G;
  for (int k = 0; k < 16; ++k) G.g[k] = make_float2(static_cast<float>(g[k].real()), static_cast<float>(g[k].imag()));
  hipLaunchKernelGGL(k_mps_merge, dim3(blocks_for(static_cast<uint64_t>(dl) * dr)), dim3(kThreads), 0, m->stream,
                     site_ptr(m, i), site_ptr(m, i + 1), m->work, dl, dm, dr, G);
  MPS_HIP(hipGetLastError());
  const int rows = 2 * dl, cols = 2 * dr, kmin = std::min(rows, cols);
  MPS_TRY(svd(m, rows, cols));
  const std::vector<float> &s = m->h_sv;
  // Exact zeros are dropped even at eps = 0: the device SVD's vectors for
  // them are not reliable, and they carry no weight. At eps = 0 (where the
  // reference keeps every value) dropping them is no truncation event; at
  // eps > 0 an event is any split below the full rank, as in the reference.
  int r_nz = kmin, r_eps = kmin;
  if (kmin > 0) {
    r_nz = r_eps = 0;
    const double s0 = static_cast<double>(s[0]);
    for (int k = 0; k < kmin; ++k) {
      r_nz += static_cast<double>(s[k]) >= kZeroFloor * s0 ? 1 : 0;
      r_eps += static_cast<double>(s[k]) >= eps * s0 ? 1 : 0;
    }
    r_nz = std::max(1, r_nz);
  }
  const int r = std::min(r_nz, std::max(1, std::min(m->dmax, r_eps)));
  double total = 0, kept = 0;
  for (int k = 0; k < kmin; ++k) {
    const double w = static_cast<double>(s[k]) * s[k];
    total += w;
    if (k < r) kept += w;
  }
  if (!(kept > 1e-300) || !std::isfinite(kept))
    return set_error(MSV_E_INVALID, "MPS truncation retained a zero or non-finite singular spectrum");
  const double discarded = total - kept;
  if (r < (eps > 0.0 ? kmin : r_nz)) {
    m->st.truncated = 1;
    m->st.trunc_events += 1;
    m->st.local_discarded_weight_sum += discarded;
    m->st.local_discarded_weight_max = std::max(m->st.local_discarded_weight_max, discarded);
    const double rel = total > 0 ? discarded / total : 0.0;
    m->st.relative_discarded_weight_sum += rel;
    m->st.relative_discarded_weight_max = std::max(m->st.relative_discarded_weight_max, rel);
  }
  const float scale = m->renormalize ? static_cast<float>(1.0 / std::sqrt(kept)) : 1.0f;
  if (m->renormalize) m->st.renormalization_count += 1;
  const size_t slot = static_cast<size_t>(4) * m->dmax * m->dmax;
  hipLaunchKernelGGL(k_mps_split, dim3(blocks_for(static_cast<uint64_t>(rows) * r + static_cast<uint64_t>(r) * cols)),
                     dim3(kThreads), 0, m->stream, m->work + slot, m->sv, m->work + 2 * slot, site_ptr(m, i),
                     site_ptr(m, i + 1), dl, dr, kmin, r, scale, 0);
  MPS_HIP(hipGetLastError());
  m->dim[i + 1] = r;
  m->center = i + 1;
  m->st.two_site_calls += 1;
  m->st.max_bond_ever = std::max<int64_t>(m->st.max_bond_ever, r);
  return MSV_OK;
}

const std::complex<double> *swap_gate() {
  static const std::complex<double> g[16] = {1, 0, 0, 0, 0, 0, 1, 0, 0, 1, 0, 0, 0, 0, 0, 1};
  return g;
}

// SWAP of sites (i, i + 1); eps floored at 1e-6 when positive (mps_state.py:84, 711-742).
int swap_adjacent(msv_mps *m, int i) {
  const double e = m->eps > 0.0 ? std::max(m->eps, 1e-6) : m->eps;
  m->st.swap_calls += 1;
  return two_site(m, i, swap_gate(), e);
}

void reset_tensors_host(msv_mps *m, std::vector<C> *h) {
  h->assign(m->cap, make_float2(0, 0));
  (*h)[0] = make_float2(1, 0);  // [1, 0] as a (1, 2, 1) tensor
}

}  // namespace

extern "C" {

int msv_mps_create(int n_qubits, int dmax, double eps, int renormalize, int device, msv_mps **out) {
  if (!out) return set_error(MSV_E_INVALID, "null output");
  *out = nullptr;
  if (n_qubits < 1 || n_qubits > 4096) return set_error(MSV_E_INVALID, "MPS qubit count out of range");
  if (dmax < 1 || dmax > 1024) return set_error(MSV_E_INVALID, "MPS dmax must be in [1, 1024]");
  // the Jacobi SVD keeps one float and one int per column of the
  // (2 dmax)-column two-site matrix in LDS
  if (!(eps >= 0.0)) return set_error(MSV_E_INVALID, "MPS eps must be non-negative");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
    return set_error(MSV_E_INVALID, "no such HIP device");
  MPS_HIP(hipSetDevice(device));
  auto *m = new msv_mps();
  m->n = n_qubits;
  m->dmax = dmax;
  m->eps = eps;
  m->renormalize = renormalize != 0;
  m->device = device;
  m->cap = static_cast<size_t>(dmax) * 2 * dmax;
  const size_t slot = static_cast<size_t>(4) * dmax * dmax;
  auto cleanup = [&](int rc) {
    msv_mps_destroy(m);
    return rc;
  };
  if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess)
    return cleanup(set_error(MSV_E_HIP, "stream creation failed"));
  if (hipMalloc(reinterpret_cast<void **>(&m->sites), sizeof(C) * m->cap * n_qubits) != hipSuccess ||
      hipMalloc(reinterpret_cast<void **>(&m->spare), sizeof(C) * m->cap) != hipSuccess ||
      hipMalloc(reinterpret_cast<void **>(&m->work), sizeof(C) * slot * 5) != hipSuccess ||
      hipMalloc(reinterpret_cast<void **>(&m->sv), sizeof(float) * 4 * dmax) != hipSuccess ||
      hipMalloc(reinterpret_cast<void **>(&m->info), sizeof(int)) != hipSuccess)
    return cleanup(set_error(MSV_E_OOM, "MPS device allocation failed"));
  const int rc = msv_mps_reset(m);
  if (rc != MSV_OK) return cleanup(rc);
  *out = m;
  return MSV_OK;
}

int msv_mps_destroy(msv_mps *m) {
  if (!m) return MSV_OK;
  hipSetDevice(m->device);
  if (m->stream) hipStreamSynchronize(m->stream);
  hipFree(m->sites);
  hipFree(m->spare);
  hipFree(m->work);
  hipFree(m->sv);
  hipFree(m->info);
  if (m->stream) hipStreamDestroy(m->stream);
  delete m;
  return MSV_OK;
}

int msv_mps_reset(msv_mps *m) {
  MPS_TRY(check_mps(m));
  MPS_HIP(hipSetDevice(m->device));
  std::vector<C> h;
  reset_tensors_host(m, &h);
  for (int i = 0; i < m->n; ++i)
    MPS_HIP(hipMemcpyAsync(site_ptr(m, i), h.data(), sizeof(C) * 2, hipMemcpyHostToDevice, m->stream));
  MPS_HIP(hipStreamSynchronize(m->stream));
  m->dim.assign(m->n + 1, 1);
  m->center = 0;
  std::memset(&m->st, 0, sizeof(m->st));
  m->st.max_bond_ever = 1;
  return MSV_OK;
}

int msv_mps_apply_1q(msv_mps *m, int q, const double *u) {
  MPS_TRY(check_mps(m));
  MPS_TRY(check_site(m, q));
  if (!u) return set_error(MSV_E_INVALID, "null gate");
  MPS_HIP(hipSetDevice(m->device));
  Gate2 U;
  for (int k = 0; k < 4; ++k) U.u[k] = make_float2(static_cast<float>(u[2 * k]), static_cast<float>(u[2 * k + 1]));
  const int dl = m->dim[q], dr = m->dim[q + 1];
  hipLaunchKernelGGL(k_mps_1q, dim3(blocks_for(static_cast<uint64_t>(dl) * dr)), dim3(kThreads), 0, m->stream,
                     site_ptr(m, q), dl, dr, U);
  MPS_HIP(hipGetLastError());
  return MSV_OK;
}

int msv_mps_apply_2q(msv_mps *m, int q0, int q1, const double *u) {
  MPS_TRY(check_mps(m));
  MPS_TRY(check_site(m, q0));
  MPS_TRY(check_site(m, q1));
  if (q0 == q1) return set_error(MSV_E_INVALID, "MPS two-qubit gate operands must differ");
  if (!u) return set_error(MSV_E_INVALID, "null gate");
  MPS_HIP(hipSetDevice(m->device));
  std::complex<double> g[16];
  for (int k = 0; k < 16; ++k) g[k] = std::complex<double>(u[2 * k], u[2 * k + 1]);
  // descending operands: conjugate by SWAP so the tensor's (left, right)
  // legs are (q0, q1) in ascending site order (mps_state.py:757-763)
  if (q0 > q1) {
    std::complex<double> h[16];
    auto sw = [](int x) { return ((x & 1) << 1) | (x >> 1); };
    for (int r = 0; r < 4; ++r)
      for (int c = 0; c < 4; ++c) h[r * 4 + c] = g[sw(r) * 4 + sw(c)];
    std::memcpy(g, h, sizeof(g));
  }
  const int i = std::min(q0, q1), j = std::max(q0, q1);
  // swap network: the lower site walks up to j - 1, then back (mps_state.py:744-768)
  for (int k = i; k < j - 1; ++k) MPS_TRY(swap_adjacent(m, k));
  MPS_TRY(two_site(m, j - 1, g, m->eps));
  for (int k = j - 2; k >= i; --k) MPS_TRY(swap_adjacent(m, k));
  return MSV_OK;
}

int msv_mps_bond_dims(msv_mps *m, int32_t *out) {
  MPS_TRY(check_mps(m));
  if (!out) return set_error(MSV_E_INVALID, "null output");
  for (int i = 0; i + 1 < m->n; ++i) out[i] = m->dim[i + 1];
  return MSV_OK;
}

int msv_mps_get_stats(msv_mps *m, msv_mps_stats *out) {
  MPS_TRY(check_mps(m));
  if (!out) return set_error(MSV_E_INVALID, "null output");
  *out = m->st;
  out->center = m->center;
  return MSV_OK;
}

int msv_mps_sync(msv_mps *m) {
  MPS_TRY(check_mps(m));
  MPS_HIP(hipSetDevice(m->device));
  MPS_HIP(hipStreamSynchronize(m->stream));
  return MSV_OK;
}

int msv_mps_site_tensor(msv_mps *m, int q, void *host, int32_t *dl, int32_t *dr) {
  MPS_TRY(check_mps(m));
  MPS_TRY(check_site(m, q));
  MPS_HIP(hipSetDevice(m->device));
  if (dl) *dl = m->dim[q];
  if (dr) *dr = m->dim[q + 1];
  if (host) {
    MPS_HIP(hipMemcpyAsync(host, site_ptr(m, q), sizeof(C) * m->dim[q] * 2 * m->dim[q + 1], hipMemcpyDeviceToHost,
                           m->stream));
    MPS_HIP(hipStreamSynchronize(m->stream));
  }
  return MSV_OK;
}

int msv_mps_norm(msv_mps *m, double *out) {
  MPS_TRY(check_mps(m));
  if (!out) return set_error(MSV_E_INVALID, "null output");
  MPS_HIP(hipSetDevice(m->device));
  const int c = m->center;
  const size_t len = static_cast<size_t>(m->dim[c]) * 2 * m->dim[c + 1];
  std::vector<C> h(len);
  MPS_HIP(hipMemcpyAsync(h.data(), site_ptr(m, c), sizeof(C) * len, hipMemcpyDeviceToHost, m->stream));
  MPS_HIP(hipStreamSynchronize(m->stream));
  double s = 0;
  for (const C &v : h) s += static_cast<double>(v.x) * v.x + static_cast<double>(v.y) * v.y;
  *out = std::sqrt(s);
  return MSV_OK;
}

int msv_mps_expval_product(msv_mps *m, const int32_t *wires, const double *ops, int k, double *re, double *im) {
  MPS_TRY(check_mps(m));
  if (!re || !im || (k > 0 && (!wires || !ops))) return set_error(MSV_E_INVALID, "null argument");
  MPS_HIP(hipSetDevice(m->device));
  std::vector<int> which(m->n, -1);
  for (int i = 0; i < k; ++i) {
    MPS_TRY(check_site(m, wires[i]));
    if (which[wires[i]] >= 0) return set_error(MSV_E_INVALID, "duplicate qubit in the product operator");
    which[wires[i]] = i;
  }
  double2 *env = nullptr;
  const size_t emax = static_cast<size_t>(m->dmax) * m->dmax;
  MPS_HIP(hipMallocAsync(reinterpret_cast<void **>(&env), sizeof(double2) * 2 * emax, m->stream));
  const double2 one = make_double2(1.0, 0.0);
  hipError_t e = hipMemcpyAsync(env, &one, sizeof(double2), hipMemcpyHostToDevice, m->stream);
  int cur = 0;
  for (int i = 0; i < m->n && e == hipSuccess; ++i) {
    Gate2 O;
    if (which[i] >= 0) {
      const double *o = ops + 8 * which[i];
      for (int t = 0; t < 4; ++t) O.u[t] = make_float2(static_cast<float>(o[2 * t]), static_cast<float>(o[2 * t + 1]));
    } else {
      O.u[0] = O.u[3] = make_float2(1, 0);
      O.u[1] = O.u[2] = make_float2(0, 0);
    }
    const int dm = m->dim[i], dn = m->dim[i + 1];
    hipLaunchKernelGGL(k_mps_transfer, dim3(blocks_for(static_cast<uint64_t>(dn) * dn)), dim3(kThreads), 0,
                       m->stream, env + cur * emax, site_ptr(m, i), env + (1 - cur) * emax, dm, dn, O);
    e = hipGetLastError();
    cur = 1 - cur;
  }
  double2 h = make_double2(0, 0);
  if (e == hipSuccess) e = hipMemcpyAsync(&h, env + cur * emax, sizeof(double2), hipMemcpyDeviceToHost, m->stream);
  hipFreeAsync(env, m->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
  if (e != hipSuccess) return fail_hip(e, "MPS expectation");
  double nrm = 0;
  MPS_TRY(msv_mps_norm(m, &nrm));
  if (!(nrm > 0)) return set_error(MSV_E_INVALID, "MPS with zero norm");
  *re = h.x / (nrm * nrm);
  *im = h.y / (nrm * nrm);
  return MSV_OK;
}

int msv_mps_to_state(msv_mps *m, msv_state *dst) {
  MPS_TRY(check_mps(m));
  if (!dst) return set_error(MSV_E_INVALID, "null state");
  int nq = 0, dtype = 0;
  uint64_t bytes = 0;
  MPS_TRY(msv_state_info(dst, &nq, &dtype, &bytes));
  if (nq != m->n) return set_error(MSV_E_INVALID, "state and MPS qubit counts differ");
  if (dtype != MSV_C64) return set_error(MSV_E_INVALID, "the MPS contracts into a complex64 state");
  MPS_TRY(msv_state_sync(dst));
  void *buf = msv_state_device_ptr(dst);
  if (!buf) return set_error(MSV_E_INVALID, "state has no device buffer");
  MPS_HIP(hipSetDevice(m->device));
  double nrm = 0;
  MPS_TRY(msv_mps_norm(m, &nrm));
  if (!(nrm > 0)) return set_error(MSV_E_INVALID, "MPS produced a zero statevector");
  const uint64_t amps = 1ull << m->n;
  C *tmp = nullptr;
  if (m->n > 1) MPS_HIP(hipMallocAsync(reinterpret_cast<void **>(&tmp), sizeof(C) * amps, m->stream));
  // ping-pong so that the last step lands in the state's buffer
  C *bufs[2] = {reinterpret_cast<C *>(buf), tmp};
  int out = (m->n - 1) % 2 == 0 ? 0 : 1;  // buffer of step 0
  hipError_t e = hipMemcpyAsync(bufs[out], site_ptr(m, 0), sizeof(C) * 2 * m->dim[1], hipMemcpyDeviceToDevice,
                                m->stream);
  uint64_t np = 2;
  for (int i = 1; i < m->n && e == hipSuccess; ++i) {
    const int dm = m->dim[i], dn = m->dim[i + 1];
    const float scale = (i == m->n - 1) ? static_cast<float>(1.0 / nrm) : 1.0f;
    hipLaunchKernelGGL(k_mps_contract, dim3(blocks_for(np * 2 * dn)), dim3(kThreads), 0, m->stream, bufs[out],
                       site_ptr(m, i), bufs[1 - out], np, dm, dn, scale);
    e = hipGetLastError();
    out = 1 - out;
    np *= 2;
  }
  if (m->n == 1 && e == hipSuccess) {
    // one qubit: normalize in place through a host round trip (two amplitudes)
    C h[2];
    e = hipMemcpyAsync(h, buf, sizeof(h), hipMemcpyDeviceToHost, m->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    for (C &v : h) {
      v.x = static_cast<float>(v.x / nrm);
      v.y = static_cast<float>(v.y / nrm);
    }
    if (e == hipSuccess) e = hipMemcpyAsync(buf, h, sizeof(h), hipMemcpyHostToDevice, m->stream);
  }
  if (tmp) hipFreeAsync(tmp, m->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
  if (e != hipSuccess) return fail_hip(e, "MPS contraction");
  return MSV_OK;
}

}  // extern "C"
