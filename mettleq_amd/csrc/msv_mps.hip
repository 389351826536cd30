// Matrix-product-state engine on one MI355X (gfx950): the reference's
// MPSState (mps_state.py:250-1549, native core _mps_native.cpp) with every
// site tensor resident in HBM. Tensor A[i] has shape (dl, 2, dr), stored
// row-major (index (a * 2 + s) * dr + b), complex64 like the reference's
// storage. Site i is qubit i (qubit 0 = the most significant amplitude bit).
//
// Every factorization here is "orthonormal basis Q of one side, then the
// projection X = Q^H M": for M = U S Vh, Q = U (the columns kept) and
// X = S Vh, so no right-hand vectors are ever accumulated.
// Per two-site update (sites i, i + 1):
//   k_mps_merge     theta = G . (A_i A_{i+1}), written column-major as the
//                   (2 dl) x (2 dr) matrix M;
//   k_mps_basis_lds one-sided Jacobi in one workgroup with the matrix in LDS
//                   (both sides up to 128; k_mps_svd, in global memory,
//                   beyond): Q = U sorted by singular value, and S (the
//                   reference: LAPACK cgesdd on the CPU). With both sides
//                   >= 32, M is first factored by k_mps_qrcp_lds (pivoted
//                   QR, M P = Q_p R), then R^T by the same kernel without
//                   pivoting (R^T = Q_2 R_2, R_2 only); the Jacobi runs on
//                   L = R_2^T, whose normalized columns are U_R, the left
//                   singular vectors of R (R = L Q_2^T), and k_mps_hh_apply
//                   forms U = Q_p U_R: a few sweeps instead of many on
//                   graded spectra, with no rotations accumulated;
//   host            the reference's truncation rule on the singular values
//                   (mps_state.py:193-247), read back (min(2 dl, 2 dr) floats);
//   k_mps_project   X = Q^H M;
//   k_mps_split     A_i = Q[:, :r], A_{i+1} = X[:r, :] / norm.
// The orthogonality center moves like the reference's QR moves
// (mps_state.py:347-399), without truncation: right, Q of the site as a
// (2 dl) x dr matrix and X carried into the next site; left, the same on
// the conjugate transpose of the site as a dl x (2 dr) matrix
// (k_mps_absorb_*).
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
// column-major one: M = A (rows x cols), or M = A^H (cols x rows) when herm.
__global__ void __launch_bounds__(256) k_mps_to_colmajor(const C *__restrict__ A, C *__restrict__ M, int rows,
                                                         int cols, int herm) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= rows * cols) return;
  const int r = idx / cols, c = idx % cols;
  const C a = A[idx];
  if (herm)
    M[c + static_cast<size_t>(r) * cols] = make_float2(a.x, -a.y);
  else
    M[r + static_cast<size_t>(c) * rows] = a;
}

// X = Q^H M (kmin x n, column-major, leading dimension kmin) for Q (m x kmin)
// and M (m x n), both column-major; sums in double precision. With M = U S Vh
// and Q = U this is S Vh.
__global__ void __launch_bounds__(256) k_mps_project(const C *__restrict__ Q, const C *__restrict__ M,
                                                     C *__restrict__ X, int m, int n, int kmin) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= kmin * n) return;
  const int j = idx % kmin, c = idx / kmin;
  const C *q = Q + static_cast<size_t>(j) * m;
  const C *a = M + static_cast<size_t>(c) * m;
  double re = 0, im = 0;
  for (int i = 0; i < m; ++i) {
    const C u = q[i], v = a[i];
    re += (double)u.x * v.x + (double)u.y * v.y;
    im += (double)u.x * v.y - (double)u.y * v.x;
  }
  X[idx] = make_float2((float)re, (float)im);
}

// Split after a two-site factorization of M = theta ((2 dl) x (2 dr)):
// left (dl, 2, r) = Q[:, :r], right (r, 2, dr) = scale X[:r, :], with Q
// (2 dl x kmin) and X (kmin x 2 dr, leading dimension kmin).
__global__ void __launch_bounds__(256) k_mps_split(const C *__restrict__ Q, const C *__restrict__ X,
                                                   C *__restrict__ left, C *__restrict__ right, int dl, int dr,
                                                   int kmin, int r, float scale) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  const int m = 2 * dl, n = 2 * dr;
  const int nl = m * r, nr = r * n;
  if (idx < nl) {
    const int row = idx / r, j = idx % r;  // row = a * 2 + s
    left[idx] = Q[row + static_cast<size_t>(j) * m];
  } else if (idx < nl + nr) {
    const int k = idx - nl;
    const int j = k / n, col = k % n;  // col = t * dr + b; right index (j * 2 + t) * dr + b = j * n + col
    C v = X[j + static_cast<size_t>(col) * kmin];
    v.x *= scale;
    v.y *= scale;
    right[k] = v;
  }
}

// Center moves right: the site as M ((2 dl) x dr) = Q X; A_i <- Q
// ((2 dl) x kmin) and the next site absorbs X (kmin x dr):
// next'[j, t, b] = sum_c X[j, c] next[c, t, b].
__global__ void __launch_bounds__(256) k_mps_absorb_right(const C *__restrict__ Q, const C *__restrict__ X,
                                                          C *__restrict__ site, const C *__restrict__ next,
                                                          C *__restrict__ next_out, int rows, int cols, int kmin,
                                                          int dn) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  const int nu = rows * kmin, nn = kmin * 2 * dn;
  if (idx < nu) {
    const int row = idx / kmin, j = idx % kmin;
    site[idx] = Q[row + static_cast<size_t>(j) * rows];
  } else if (idx < nu + nn) {
    const int k = idx - nu;
    const int j = k / (2 * dn), tb = k % (2 * dn);
    C acc = make_float2(0, 0);
    for (int c = 0; c < cols; ++c) acc = cadd(acc, cmul(X[j + static_cast<size_t>(c) * kmin], next[c * 2 * dn + tb]));
    next_out[k] = acc;
  }
}

// Center moves left: the site (dl x (2 dr)) enters as M = site^H
// ((2 dr) x dl) = Q X, so site = X^H Q^H; A_i <- Q^H (kmin x (2 dr)) and the
// previous site absorbs X^H (dl x kmin): prev'[a, t, j] = sum_c prev[a, t, c] conj(X[j, c]).
__global__ void __launch_bounds__(256) k_mps_absorb_left(const C *__restrict__ Q, const C *__restrict__ X,
                                                         C *__restrict__ site, const C *__restrict__ prev,
                                                         C *__restrict__ prev_out, int dl, int cols, int kmin,
                                                         int dp) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  const int nv = kmin * cols, np = dp * 2 * kmin;  // cols = 2 dr, the rows of Q
  if (idx < nv) {
    const int j = idx / cols, c = idx % cols;
    const C q = Q[c + static_cast<size_t>(j) * cols];
    site[idx] = make_float2(q.x, -q.y);
  } else if (idx < nv + np) {
    const int k = idx - nv;
    const int at = k / kmin, j = k % kmin;  // at = a * 2 + t
    C acc = make_float2(0, 0);
    for (int c = 0; c < dl; ++c) {
      const C x = X[j + static_cast<size_t>(c) * kmin];
      acc = cadd(acc, cmul(prev[at * dl + c], make_float2(x.x, -x.y)));
    }
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

// Transfer step of <psi| O_1 x ... x O_n |psi> (complex128 environment) in
// two contractions of O(d^3) each:
//   k_mps_transfer_a  T[a, t, b'] = sum_a' env[a, a'] A[a', t, b'];
//   k_mps_transfer_b  env'[b, b'] = sum_{a, s} conj(A[a, s, b]) sum_t O[s, t] T[a, t, b'].
__global__ void __launch_bounds__(256) k_mps_transfer_a(const double2 *__restrict__ env, const C *__restrict__ A,
                                                        double2 *__restrict__ T, int dm, int dn) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= dm * 2 * dn) return;
  const int a = idx / (2 * dn), tb = idx % (2 * dn);
  double re = 0, im = 0;
  for (int ap = 0; ap < dm; ++ap) {
    const double2 e = env[a * dm + ap];
    const C x = A[ap * 2 * dn + tb];
    re += e.x * x.x - e.y * x.y;
    im += e.x * x.y + e.y * x.x;
  }
  T[idx] = make_double2(re, im);
}

__global__ void __launch_bounds__(256) k_mps_transfer_b(const double2 *__restrict__ T, const C *__restrict__ A,
                                                        double2 *__restrict__ out, int dm, int dn, Gate2 O) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= dn * dn) return;
  const int b = idx / dn, bp = idx % dn;
  double re = 0, im = 0;
  for (int a = 0; a < dm; ++a) {
    const double2 t0 = T[(a * 2) * dn + bp], t1 = T[(a * 2 + 1) * dn + bp];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const C o0 = O.u[s * 2], o1 = O.u[s * 2 + 1];
      // y = O[s, 0] T[a, 0, b'] + O[s, 1] T[a, 1, b']
      const double yr = o0.x * t0.x - o0.y * t0.y + o1.x * t1.x - o1.y * t1.y;
      const double yi = o0.x * t0.y + o0.y * t0.x + o1.x * t1.y + o1.y * t1.x;
      const C ca = A[(a * 2 + s) * dn + b];  // conj taken here
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
  // outputs: the left vectors of W are w_j / sigma_j, its right vectors V;
  // columns at or below the floor never rotated and stay zero
  for (int idx = tid; idx < kk * k; idx += blockDim.x) {
    const int i = idx % kk, j = idx / kk;
    const float inv = (double)sig[j] * sig[j] > floor2 ? 1.f / sig[j] : 0.f;
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

// Fast path of the factorization for m, n <= kBasisMaxDim: the same
// one-sided Jacobi (same pair order, rotation, tolerance and floor as
// k_mps_svd) with the working matrix in LDS. Each 16-lane DPP row owns one
// column pair per round and holds 8 rows of each column in registers; the
// four dot products reduce over the row by DPP (all 16 lanes end with the
// same bits, so the rotate/skip decision is uniform in the row).
//   !kAccum: W = M (m x n); the columns of W converge to U S, so
//            Q[:, rank j] = w_j / |w_j| (the first kmin of them);
//   kAccum:  W = M^H (n x m), with the rotations accumulated in V (m x m;
//            in LDS beside W);
//            W V = V_M S, so Q[:, rank j] = v_j.
// S gets the kmin largest column norms in descending order; info = sweeps
// run, or -1 without convergence.
constexpr int kBasisMaxDim = 128;
constexpr int kBasisLds = 16384;  // complex64 entries of W (+ V): 128 KiB
constexpr int kPairLanes = 16;
constexpr int kRowsPerLane = kBasisMaxDim / kPairLanes;

template <int Ctrl>
__device__ __forceinline__ double dpp_f64(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), Ctrl, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), Ctrl, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

// Sum over the 16 lanes of a DPP row: lane pairs (i, i ^ 1), (i, i ^ 2), then
// the mirrored half-rows and halves, so each add sees the same two operands
// in both lanes of the pair.
__device__ __forceinline__ double row16_sum(double v) {
  v += dpp_f64<0xB1>(v);   // quad_perm [1, 0, 3, 2]
  v += dpp_f64<0x4E>(v);   // quad_perm [2, 3, 0, 1]
  v += dpp_f64<0x141>(v);  // row_half_mirror
  v += dpp_f64<0x140>(v);  // row_mirror
  return v;
}

// 1 / sqrt(x) and 1 / x in double from the hardware approximations plus
// one Newton step (relative error far below float rounding).
__device__ __forceinline__ double rsq_nr(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  return y * (1.5 - 0.5 * x * y * y);
}
__device__ __forceinline__ double rcp_nr(double x) {
  const double y = __builtin_amdgcn_rcp(x);
  return y * (2.0 - x * y);
}

template <bool kAccum, int RPL>
__global__ void __launch_bounds__(1024) k_mps_basis_lds(const C *__restrict__ M, int m, int n, C *__restrict__ Q,
                                                        float *__restrict__ S, int *__restrict__ info) {
  static_assert(RPL >= 1 && RPL <= kRowsPerLane, "row slots per lane");
  __shared__ C lds[kBasisLds];
  __shared__ float sig[kBasisMaxDim];
  __shared__ int rank[kBasisMaxDim];
  __shared__ double wave_fro[kSvdThreads / 64];
  __shared__ int rotated;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nwave = blockDim.x >> 6;
  const int l16 = tid & (kPairLanes - 1), group = tid / kPairLanes, ngroups = blockDim.x / kPairLanes;
  const int kk = kAccum ? n : m;  // rows of W
  const int k = kAccum ? m : n;   // columns of W
  C *W = lds;
  C *V = lds + kk * k;  // the rotations' accumulation (kAccum)
  double fro = 0;
  for (int idx = tid; idx < kk * k; idx += blockDim.x) {
    C a;
    if (!kAccum) {
      a = M[idx];
    } else {
      const int i = idx % kk, j = idx / kk;
      a = M[j + static_cast<size_t>(i) * m];  // W[i, j] = conj(M[j, i])
      a.y = -a.y;
    }
    W[idx] = a;
    fro += (double)a.x * a.x + (double)a.y * a.y;
  }
  if (kAccum)
    for (int idx = tid; idx < k * k; idx += blockDim.x)
      V[idx] = make_float2((idx % k) == (idx / k) ? 1.f : 0.f, 0.f);
  fro = wave_sum(fro);
  if (lane == 0) wave_fro[wave] = fro;
  __syncthreads();
  fro = 0;
  for (int w = 0; w < nwave; ++w) fro += wave_fro[w];
  const double floor2 = fmax(kJacobiFloor * kJacobiFloor * fro, 1e-66);
  const int np = k + (k & 1), npairs = np / 2;
  const int pi = group;  // one pair slot per 16-lane row (the launch has npairs rows or more)
  int sweeps = 0;
  bool converged = k < 2;
  for (; sweeps < kJacobiMaxSweeps && !converged; ++sweeps) {
    if (tid == 0) rotated = 0;
    __syncthreads();
    // round-robin positions at round 0: a = pi (0 stays fixed), b = np - 1 - pi; both step
    // through 1 .. np - 1 cyclically from round to round
    int a = pi, b = np - 1 - pi;
    for (int r = 0; r < np - 1; ++r) {
      if (pi < npairs) {
        const int p = a < b ? a : b, q = a < b ? b : a;
        if (q < k) {
          C *wp = W + p * kk, *wq = W + q * kk;
          // RPL * 16 >= kk row slots; slots past kk read row 0 and count as zero,
          // so every read is issued before the first use
          C x[RPL], y[RPL];
#pragma unroll
          for (int t = 0; t < RPL; ++t) {
            const int i = l16 + kPairLanes * t;
            const int ii = i < kk ? i : 0;
            x[t] = wp[ii];
            y[t] = wq[ii];
          }
          double al = 0, be = 0, gr = 0, gi = 0;
#pragma unroll
          for (int t = 0; t < RPL; ++t) {
            if (l16 + kPairLanes * t >= kk) x[t] = y[t] = make_float2(0.f, 0.f);
            al += (double)x[t].x * x[t].x + (double)x[t].y * x[t].y;
            be += (double)y[t].x * y[t].x + (double)y[t].y * y[t].y;
            gr += (double)x[t].x * y[t].x + (double)x[t].y * y[t].y;  // conj(x) y
            gi += (double)x[t].x * y[t].y - (double)x[t].y * y[t].x;
          }
          al = row16_sum(al);
          be = row16_sum(be);
          gr = row16_sum(gr);
          gi = row16_sum(gi);
          const double g2 = gr * gr + gi * gi;
          if (g2 > kJacobiTol * kJacobiTol * (al * be) && fmin(al, be) > floor2) {
            // t = sign(z) / (|z| + sqrt(1 + z^2)), z = (be - al) / (2 g), and
            // e^{-i phi} = conj(gamma) / g, to float accuracy (they rotate
            // float data): hardware reciprocal and reciprocal square root
            // with one Newton step each
            const double ig = rsq_nr(g2);
            const double z = 0.5 * (be - al) * ig;
            const double w = 1.0 + z * z;
            const double tn = rcp_nr(fabs(z) + w * rsq_nr(w));
            // c and s in double, each rounded once: c^2 + s^2 stays within
            // float rounding of 1, which the column norms accumulate over
            // every rotation they take part in
            const double cd = rsq_nr(1.0 + tn * tn);
            const float fc = (float)cd, fs = (float)(z >= 0 ? cd * tn : -cd * tn);
            const float fr = (float)(gr * ig), fi = (float)(-gi * ig);
#pragma unroll
            for (int t = 0; t < RPL; ++t) {
              const int i = l16 + kPairLanes * t;
              if (i < kk) {
                const C ey = make_float2(fr * y[t].x - fi * y[t].y, fr * y[t].y + fi * y[t].x);
                wp[i] = make_float2(fc * x[t].x - fs * ey.x, fc * x[t].y - fs * ey.y);
                wq[i] = make_float2(fs * x[t].x + fc * ey.x, fs * x[t].y + fc * ey.y);
              }
            }
            if (kAccum) {
              // V has k = m < n = kk rows, so RPL slots cover them too
              C *vp = V + p * k, *vq = V + q * k;
              C vx[RPL], vy[RPL];
#pragma unroll
              for (int t = 0; t < RPL; ++t) {
                const int i = l16 + kPairLanes * t;
                const int ii = i < k ? i : 0;
                vx[t] = vp[ii];
                vy[t] = vq[ii];
              }
#pragma unroll
              for (int t = 0; t < RPL; ++t) {
                const int i = l16 + kPairLanes * t;
                if (i < k) {
                  const C xv = vx[t], yv = vy[t];
                  const C ey = make_float2(fr * yv.x - fi * yv.y, fr * yv.y + fi * yv.x);
                  vp[i] = make_float2(fc * xv.x - fs * ey.x, fc * xv.y - fs * ey.y);
                  vq[i] = make_float2(fs * xv.x + fc * ey.x, fs * xv.y + fc * ey.y);
                }
              }
            }
            if (l16 == 0) rotated = 1;
          }
        }
        if (pi != 0) a = a == np - 1 ? 1 : a + 1;
        b = b == np - 1 ? 1 : b + 1;
      }
      __syncthreads();
    }
    converged = rotated == 0;
    __syncthreads();
  }
  for (int j = group; j < k; j += ngroups) {
    double s2 = 0;
#pragma unroll
    for (int t = 0; t < RPL; ++t) {
      const int i = l16 + kPairLanes * t;
      if (i < kk) {
        const C x = W[j * kk + i];
        s2 += (double)x.x * x.x + (double)x.y * x.y;
      }
    }
    s2 = row16_sum(s2);
    if (l16 == 0) sig[j] = (float)sqrt(s2);
  }
  __syncthreads();
  const int kmin = m < n ? m : n;
  for (int j = tid; j < k; j += blockDim.x) {
    int rk = 0;
    for (int l = 0; l < k; ++l) rk += (sig[l] > sig[j] || (sig[l] == sig[j] && l < j)) ? 1 : 0;
    rank[j] = rk;
    if (rk < kmin) S[rk] = sig[j];
  }
  __syncthreads();
  if (!kAccum) {
    // columns at or below the floor never rotated: their directions are not
    // orthogonal to the others, so they stay zero in Q (weight <= 1e-32 |M|)
    for (int idx = tid; idx < kk * k; idx += blockDim.x) {
      const int i = idx % kk, j = idx / kk, rk = rank[j];
      if (rk < kmin) {
        const float inv = (double)sig[j] * sig[j] > floor2 ? 1.f / sig[j] : 0.f;
        const C w = W[idx];
        Q[i + static_cast<size_t>(rk) * m] = make_float2(w.x * inv, w.y * inv);
      }
    }
  } else {
    for (int idx = tid; idx < k * k; idx += blockDim.x) {
      const int i = idx % k, j = idx / k;
      Q[i + static_cast<size_t>(rank[j]) * m] = V[idx];
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
  C *sites = nullptr;             // the n site buffers' allocation
  C *spare_block = nullptr;       // one more site buffer's allocation
  std::vector<C *> site;          // buffer of each site: a center move absorbs into the spare
  C *spare = nullptr;             // buffer, then trades places with the site's (no copy)
  C *work = nullptr;              // 5 slots of 4 dmax^2: the factorization's input M, Q, X = Q^H M,
                                  // and W and V of the global-memory Jacobi (k_mps_svd)
  float *sv = nullptr;            // singular values (up to 2 dmax)
  double *hh_beta = nullptr;      // reflector factors of the pivoted QR (up to 2 dmax)
  int *hh_piv = nullptr;          // its pivots (up to 2 dmax)
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
C *site_ptr(msv_mps *m, int i) { return m->site[i]; }

// Reduced QR of M (rows x cols, column-major; rows * cols + rows * kmin <=
// kBasisLds) in one workgroup with M in LDS, for the center moves (the
// reference's economic QR, mps_state.py:347-399): Householder reflectors
// H_j = I - beta_j u_j u_j^H (Hermitian and unitary) with
// u_j = x + e^{i arg x_0} |x| e_0 for x = M[j:, j], so H_j x = -e^{i arg x_0} |x| e_0;
// dot products in double precision. Q = H_0 ... H_{kmin-1} [I; 0] (rows x kmin)
// is orthonormal whatever the rank of M; R = Q^H M (kmin x cols, leading
// dimension kmin, upper trapezoidal). A column whose x is at or below the
// floor (1e-32 |M|_F) takes no reflector (beta = 0).
// Column c belongs to the 16-lane DPP row c mod (rows of the launch); each
// step applies H_j to the owned columns c > j, and the owner of column j + 1
// forms the next reflector as soon as its update is done, so each step (and
// each step of the accumulation of Q) ends in one barrier.
template <int RPL>
__device__ __forceinline__ void hh_form(C *col, int j, int rows, int l16, double floor2, double *beta,
                                        C *rdiag) {
  double s2 = 0;
  C x0 = make_float2(0.f, 0.f);
#pragma unroll
  for (int t = 0; t < RPL; ++t) {
    const int r = j + l16 + kPairLanes * t;
    if (r < rows) {
      const C x = col[r];
      s2 += (double)x.x * x.x + (double)x.y * x.y;
    }
  }
  s2 = row16_sum(s2);
  x0 = col[j];
  if (!(s2 > floor2)) {
    if (l16 == 0) {
      beta[j] = 0.0;
      rdiag[j] = x0;
    }
    return;
  }
  const double nx = sqrt(s2);
  const double a0 = sqrt((double)x0.x * x0.x + (double)x0.y * x0.y);
  // e^{i arg x_0} (1 when x_0 = 0)
  const double er = a0 > 0 ? x0.x / a0 : 1.0, ei = a0 > 0 ? x0.y / a0 : 0.0;
  if (l16 == 0) {
    col[j] = make_float2((float)(x0.x + er * nx), (float)(x0.y + ei * nx));  // u_0
    rdiag[j] = make_float2((float)(-er * nx), (float)(-ei * nx));
    beta[j] = 1.0 / (nx * (nx + a0));  // 2 / (u^H u), u^H u = 2 |x| (|x| + |x_0|)
  }
}

// col[j:] <- H col[j:] for the reflector u stored in ucol[j:].
template <int RPL>
__device__ __forceinline__ void hh_apply(const C *ucol, C *col, int j, int rows, int l16, double beta) {
  C u[RPL], x[RPL];
#pragma unroll
  for (int t = 0; t < RPL; ++t) {
    const int r = j + l16 + kPairLanes * t;
    const int rr = r < rows ? r : j;
    u[t] = ucol[rr];
    x[t] = col[rr];
  }
  double wr = 0, wi = 0;
#pragma unroll
  for (int t = 0; t < RPL; ++t) {
    if (j + l16 + kPairLanes * t < rows) {
      wr += (double)u[t].x * x[t].x + (double)u[t].y * x[t].y;  // conj(u) x
      wi += (double)u[t].x * x[t].y - (double)u[t].y * x[t].x;
    }
  }
  wr = row16_sum(wr) * beta;
  wi = row16_sum(wi) * beta;
  const float fr = (float)wr, fi = (float)wi;
#pragma unroll
  for (int t = 0; t < RPL; ++t) {
    const int r = j + l16 + kPairLanes * t;
    if (r < rows) {
      // x -= u w
      col[r] = make_float2(x[t].x - (u[t].x * fr - u[t].y * fi), x[t].y - (u[t].x * fi + u[t].y * fr));
    }
  }
}

template <int RPL>
__global__ void __launch_bounds__(1024) k_mps_qr_lds(const C *__restrict__ M, int rows, int cols,
                                                     C *__restrict__ Q, C *__restrict__ R) {
  __shared__ C lds[kBasisLds];
  __shared__ double beta[kBasisMaxDim];
  __shared__ C rdiag[kBasisMaxDim];
  __shared__ double wave_part[kSvdThreads / 64];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nwave = blockDim.x >> 6;
  const int l16 = tid & (kPairLanes - 1), group = tid / kPairLanes, ngroups = blockDim.x / kPairLanes;
  const int kmin = rows < cols ? rows : cols;
  C *A = lds;
  C *Qs = lds + rows * cols;
  double fro = 0;
  for (int idx = tid; idx < rows * cols; idx += blockDim.x) {
    const C a = M[idx];
    A[idx] = a;
    fro += (double)a.x * a.x + (double)a.y * a.y;
  }
  for (int idx = tid; idx < rows * kmin; idx += blockDim.x)
    Qs[idx] = make_float2((idx % rows) == (idx / rows) ? 1.f : 0.f, 0.f);
  fro = wave_sum(fro);
  if (lane == 0) wave_part[wave] = fro;
  __syncthreads();
  fro = 0;
  for (int w = 0; w < nwave; ++w) fro += wave_part[w];
  const double floor2 = fmax(kJacobiFloor * kJacobiFloor * fro, 1e-66);
  if (group == 0) hh_form<RPL>(A, 0, rows, l16, floor2, beta, rdiag);
  __syncthreads();
  for (int j = 0; j < kmin; ++j) {
    const double bj = beta[j];
    const C *u = A + j * rows;
    // the owner of column j + 1 updates it first, then forms the next reflector
    for (int c = j + 1 + ((group - (j + 1)) % ngroups + ngroups) % ngroups; c < cols; c += ngroups) {
      C *col = A + c * rows;
      if (bj != 0.0) hh_apply<RPL>(u, col, j, rows, l16, bj);
      if (c == j + 1 && j + 1 < kmin) hh_form<RPL>(col, j + 1, rows, l16, floor2, beta, rdiag);
    }
    __syncthreads();
  }
  // Q = H_0 ... H_{kmin-1} [I; 0]: H_j touches rows j.. and, applied in this
  // order, only the columns c >= j
  for (int j = kmin - 1; j >= 0; --j) {
    const double bj = beta[j];
    if (bj != 0.0) {
      const C *u = A + j * rows;
      for (int c = j + ((group - j) % ngroups + ngroups) % ngroups; c < kmin; c += ngroups)
        hh_apply<RPL>(u, Qs + c * rows, j, rows, l16, bj);
    }
    __syncthreads();
  }
  for (int idx = tid; idx < rows * kmin; idx += blockDim.x) Q[idx] = Qs[idx];
  for (int idx = tid; idx < kmin * cols; idx += blockDim.x) {
    const int i = idx % kmin, c = idx / kmin;
    C v = make_float2(0.f, 0.f);
    if (i == c) v = rdiag[i];
    else if (i < c) v = A[i + c * rows];
    R[idx] = v;
  }
}

// Householder QR with column pivoting (Businger-Golub) of M (rows x cols,
// column-major, rows * cols <= kBasisLds) in one workgroup, the
// preconditioner of the two-site factorization: one-sided Jacobi on the
// triangular factor of a pivoted QR (after a second, unpivoted QR of R^T)
// converges in a few sweeps where Jacobi on M itself needs many for a
// graded spectrum (Drmac and Veselic, SIAM J. Matrix Anal. Appl. 29
// (2008)). Step j takes the unpicked column of largest norm over rows j..;
// columns stay in place in LDS and piv[] names the physical column of each
// logical one. Outputs: R = Q^H M P (kmin x cols, leading dimension kmin,
// pivoted column order); the reflectors H_j = I - beta_j u_j u_j^H as A
// (u_j in rows j.. of physical column piv[j], rows x cols), beta and piv,
// for k_mps_hh_apply. Each step is two barriers: every thread finds the
// pivot from the norms of the step before and the pivot's owner forms the
// reflector; then the owners update their unpicked columns and their norms
// over rows j + 1.., and the pivot's owner marks it picked.
// kPivot = false: plain Householder QR (step j reflects column j; the norms
// are not kept), for the second factorization of the preconditioner, which
// needs R only (Aout == nullptr: no reflectors, beta or piv written).
template <int RPL>
__device__ __forceinline__ double col_norm2_from(const C *col, int r0, int rows, int l16) {
  double s2 = 0;
#pragma unroll
  for (int t = 0; t < RPL; ++t) {
    const int r = r0 + l16 + kPairLanes * t;
    if (r < rows) {
      const C x = col[r];
      s2 += (double)x.x * x.x + (double)x.y * x.y;
    }
  }
  return row16_sum(s2);
}

template <int RPL, bool kPivot = true>
__global__ void __launch_bounds__(1024) k_mps_qrcp_lds(const C *__restrict__ M, int rows, int cols,
                                                       C *__restrict__ R, C *__restrict__ Aout,
                                                       double *__restrict__ beta_out, int *__restrict__ piv_out) {
  __shared__ C A[kBasisLds];
  __shared__ double beta[kBasisMaxDim];
  __shared__ C rdiag[kBasisMaxDim];
  __shared__ double cn2[kBasisMaxDim];  // remaining norm^2 of each physical column
  __shared__ int picked[kBasisMaxDim];
  __shared__ int piv[kBasisMaxDim];     // physical column of each logical column
  __shared__ double wave_part[kSvdThreads / 64];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nwave = blockDim.x >> 6;
  const int l16 = tid & (kPairLanes - 1), group = tid / kPairLanes, ngroups = blockDim.x / kPairLanes;
  const int kmin = rows < cols ? rows : cols;
  double fro = 0;
  for (int idx = tid; idx < rows * cols; idx += blockDim.x) {
    const C a = M[idx];
    A[idx] = a;
    fro += (double)a.x * a.x + (double)a.y * a.y;
  }
  fro = wave_sum(fro);
  if (lane == 0) wave_part[wave] = fro;
  __syncthreads();
  for (int c = group; c < cols; c += ngroups) {
    const double s2 = kPivot ? col_norm2_from<RPL>(A + c * rows, 0, rows, l16) : 0.0;
    if (l16 == 0) {
      cn2[c] = s2;
      picked[c] = 0;
    }
  }
  fro = 0;
  for (int w = 0; w < nwave; ++w) fro += wave_part[w];
  const double floor2 = fmax(kJacobiFloor * kJacobiFloor * fro, 1e-66);
  __syncthreads();
  for (int j = 0; j < kmin; ++j) {
    // the pivot: argmax over the unpicked columns, lowest index on ties (a
    // NaN norm counts as 0), reduced across the wavefront; every wavefront
    // gets the same result
    int p = j;
    if (kPivot) {
      double bv = -1.0;
      int bi = 0;
      for (int c = lane; c < cols; c += 64) {
        if (picked[c]) continue;
        const double v = cn2[c] >= 0.0 ? cn2[c] : 0.0;
        if (v > bv) {
          bv = v;
          bi = c;
        }
      }
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        const double ov = __shfl_xor(bv, o, 64);
        const int oi = __shfl_xor(bi, o, 64);
        if (ov > bv || (ov == bv && oi < bi)) {
          bv = ov;
          bi = oi;
        }
      }
      p = bi;
    }
    if (p % ngroups == group) hh_form<RPL>(A + p * rows, j, rows, l16, floor2, beta, rdiag);
    __syncthreads();
    const double bj = beta[j];
    const C *u = A + p * rows;
    for (int c = group; c < cols; c += ngroups) {
      if (c == p) {
        if (l16 == 0) {
          picked[c] = 1;
          piv[j] = c;
        }
        continue;
      }
      if (picked[c]) continue;
      C *col = A + c * rows;
      if (bj != 0.0) hh_apply<RPL>(u, col, j, rows, l16, bj);
      if (kPivot) {
        const double s2 = col_norm2_from<RPL>(col, j + 1, rows, l16);
        if (l16 == 0) cn2[c] = s2;
      }
    }
    __syncthreads();
  }
  // logical columns past kmin (wide M): the unpicked ones in index order
  if (tid == 0) {
    int at = kmin;
    for (int c = 0; c < cols; ++c)
      if (!picked[c]) piv[at++] = c;
  }
  __syncthreads();
  if (Aout != nullptr) {
    for (int idx = tid; idx < rows * cols; idx += blockDim.x) Aout[idx] = A[idx];
    for (int j = tid; j < kmin; j += blockDim.x) {
      beta_out[j] = beta[j];
      piv_out[j] = piv[j];
    }
  }
  for (int idx = tid; idx < kmin * cols; idx += blockDim.x) {
    const int i = idx % kmin, c = idx / kmin;
    C v = make_float2(0.f, 0.f);
    if (i == c) v = rdiag[i];
    else if (i < c) v = A[i + piv[c] * rows];
    R[idx] = v;
  }
}

// U = H_0 ... H_{kmin-1} [Ur; 0] (rows x kmin, column-major) for the
// reflectors of k_mps_qrcp_lds and Ur (kmin x kmin): with R = Ur S Vr^H,
// M P = Q R gives M's left singular vectors U = Q Ur. U sits in LDS; each
// step stages u_j there and updates every column of U.
template <int RPL>
__global__ void __launch_bounds__(1024) k_mps_hh_apply(const C *__restrict__ A, const double *__restrict__ beta,
                                                       const int *__restrict__ piv, const C *__restrict__ Ur,
                                                       C *__restrict__ U, int rows, int kmin) {
  __shared__ C Us[kBasisLds];
  __shared__ C us[kBasisMaxDim];
  const int tid = threadIdx.x;
  const int l16 = tid & (kPairLanes - 1), group = tid / kPairLanes, ngroups = blockDim.x / kPairLanes;
  for (int idx = tid; idx < rows * kmin; idx += blockDim.x) {
    const int i = idx % rows, j = idx / rows;
    Us[idx] = i < kmin ? Ur[i + static_cast<size_t>(j) * kmin] : make_float2(0.f, 0.f);
  }
  for (int j = kmin - 1; j >= 0; --j) {
    const double bj = beta[j];
    if (bj == 0.0) continue;  // the same for every thread
    const C *ucol = A + static_cast<size_t>(piv[j]) * rows;
    __syncthreads();  // the previous step's updates (or the load) are done with us
    for (int r = j + tid; r < rows; r += blockDim.x) us[r] = ucol[r];
    __syncthreads();
    for (int c = group; c < kmin; c += ngroups) hh_apply<RPL>(us, Us + c * rows, j, rows, l16, bj);
  }
  __syncthreads();
  for (int idx = tid; idx < rows * kmin; idx += blockDim.x) U[idx] = Us[idx];
}

template <bool kAccum>
void launch_basis_rpl(int rpl, int threads, hipStream_t stream, const C *M, int m, int n, C *Q, float *S,
                      int *info) {
  switch (rpl) {
    case 1:
      hipLaunchKernelGGL((k_mps_basis_lds<kAccum, 1>), dim3(1), dim3(threads), 0, stream, M, m, n, Q, S, info);
      break;
    case 2:
      hipLaunchKernelGGL((k_mps_basis_lds<kAccum, 2>), dim3(1), dim3(threads), 0, stream, M, m, n, Q, S, info);
      break;
    case 4:
      hipLaunchKernelGGL((k_mps_basis_lds<kAccum, 4>), dim3(1), dim3(threads), 0, stream, M, m, n, Q, S, info);
      break;
    default:
      hipLaunchKernelGGL((k_mps_basis_lds<kAccum, 8>), dim3(1), dim3(threads), 0, stream, M, m, n, Q, S, info);
      break;
  }
}

int rpl_for(int rows) { return rows <= 16 ? 1 : rows <= 32 ? 2 : rows <= 64 ? 4 : 8; }

int threads_for_pairs(int k) {
  const int npairs = (k + (k & 1)) / 2;
  return std::min(kSvdThreads, std::max(64, (npairs * kPairLanes + 63) / 64 * 64));
}

void launch_basis(bool accum, int rpl, int threads, hipStream_t stream, const C *M, int m, int n, C *Q, float *S,
                  int *info) {
  if (accum)
    launch_basis_rpl<true>(rpl, threads, stream, M, m, n, Q, S, info);
  else
    launch_basis_rpl<false>(rpl, threads, stream, M, m, n, Q, S, info);
}

// Smaller side from which a two-site factorization takes the pivoted-QR
// preconditioner (below it, plain Jacobi on M is as fast).
constexpr int kPrecondMin = 32;

// Factorizes M (rows x cols, column-major in work slot 0) as M = Q X: Q
// (rows x kmin, orthonormal columns in descending singular-value order) in
// slot 1, the singular values in sv and h_sv, X = Q^H M (kmin x cols,
// leading dimension kmin) in slot 2.
int factor(msv_mps *m, int rows, int cols, bool precondition = false) {
  const size_t slot = static_cast<size_t>(4) * m->dmax * m->dmax;
  const int kmin = std::min(rows, cols);
  const bool fits_qrcp = rows <= kBasisMaxDim && cols <= kBasisMaxDim;  // M (and then U) in LDS
  if (precondition && kmin >= kPrecondMin && fits_qrcp) {
    // M P = Q R (pivoted QR), then U_R, the left singular vectors of R
    // (R = U_R S V_R^H), and U = Q U_R. Slots: the reflectors in 3, R
    // (kmin x cols) in 4, U_R in 2, U in 1.
    C *refl = m->work + 3 * slot, *r = m->work + 4 * slot, *ur = m->work + 2 * slot, *u = m->work + slot;
    const int qthreads = std::min(kSvdThreads, std::max(64, (cols * kPairLanes + 63) / 64 * 64));
    const int hthreads = std::min(kSvdThreads, std::max(64, (kmin * kPairLanes + 63) / 64 * 64));
    const int rp = rpl_for(rows);
    if (rp == 1)
      hipLaunchKernelGGL(k_mps_qrcp_lds<1>, dim3(1), dim3(qthreads), 0, m->stream, m->work, rows, cols, r, refl,
                         m->hh_beta, m->hh_piv);
    else if (rp == 2)
      hipLaunchKernelGGL(k_mps_qrcp_lds<2>, dim3(1), dim3(qthreads), 0, m->stream, m->work, rows, cols, r, refl,
                         m->hh_beta, m->hh_piv);
    else if (rp == 4)
      hipLaunchKernelGGL(k_mps_qrcp_lds<4>, dim3(1), dim3(qthreads), 0, m->stream, m->work, rows, cols, r, refl,
                         m->hh_beta, m->hh_piv);
    else
      hipLaunchKernelGGL(k_mps_qrcp_lds<8>, dim3(1), dim3(qthreads), 0, m->stream, m->work, rows, cols, r, refl,
                         m->hh_beta, m->hh_piv);
    MPS_HIP(hipGetLastError());
    // Second QR (Drmac and Veselic): R^T = Q2 R2 (R2 only) and L = R2^T, so
    // R = L Q2^T and the left singular vectors of L are U_R. Jacobi on L's
    // columns converges like Jacobi on R^H and gives them as its normalized
    // columns, with no rotations to keep (on R^H they would have to be
    // accumulated, in global memory once both exceed LDS). R^T (cols x
    // kmin) in slot 1, R2 (kmin x kmin) in 2, L in 1.
    C *rt = u, *r2 = ur, *lt = u;
    hipLaunchKernelGGL(k_mps_to_colmajor, dim3(blocks_for(static_cast<uint64_t>(kmin) * cols)), dim3(kThreads), 0,
                       m->stream, r, rt, cols, kmin, 0);
    MPS_HIP(hipGetLastError());
    const int q2threads = std::min(kSvdThreads, std::max(64, (kmin * kPairLanes + 63) / 64 * 64));
    const int rp2 = rpl_for(cols);
    if (rp2 == 1)
      hipLaunchKernelGGL((k_mps_qrcp_lds<1, false>), dim3(1), dim3(q2threads), 0, m->stream, rt, cols, kmin, r2,
                         nullptr, nullptr, nullptr);
    else if (rp2 == 2)
      hipLaunchKernelGGL((k_mps_qrcp_lds<2, false>), dim3(1), dim3(q2threads), 0, m->stream, rt, cols, kmin, r2,
                         nullptr, nullptr, nullptr);
    else if (rp2 == 4)
      hipLaunchKernelGGL((k_mps_qrcp_lds<4, false>), dim3(1), dim3(q2threads), 0, m->stream, rt, cols, kmin, r2,
                         nullptr, nullptr, nullptr);
    else
      hipLaunchKernelGGL((k_mps_qrcp_lds<8, false>), dim3(1), dim3(q2threads), 0, m->stream, rt, cols, kmin, r2,
                         nullptr, nullptr, nullptr);
    MPS_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_mps_to_colmajor, dim3(blocks_for(static_cast<uint64_t>(kmin) * kmin)), dim3(kThreads), 0,
                       m->stream, r2, lt, kmin, kmin, 0);
    MPS_HIP(hipGetLastError());
    launch_basis_rpl<false>(rpl_for(kmin), threads_for_pairs(kmin), m->stream, lt, kmin, kmin, ur, m->sv, m->info);
    MPS_HIP(hipGetLastError());
    if (rp == 1)
      hipLaunchKernelGGL(k_mps_hh_apply<1>, dim3(1), dim3(hthreads), 0, m->stream, refl, m->hh_beta, m->hh_piv, ur,
                         u, rows, kmin);
    else if (rp == 2)
      hipLaunchKernelGGL(k_mps_hh_apply<2>, dim3(1), dim3(hthreads), 0, m->stream, refl, m->hh_beta, m->hh_piv, ur,
                         u, rows, kmin);
    else if (rp == 4)
      hipLaunchKernelGGL(k_mps_hh_apply<4>, dim3(1), dim3(hthreads), 0, m->stream, refl, m->hh_beta, m->hh_piv, ur,
                         u, rows, kmin);
    else
      hipLaunchKernelGGL(k_mps_hh_apply<8>, dim3(1), dim3(hthreads), 0, m->stream, refl, m->hh_beta, m->hh_piv, ur,
                         u, rows, kmin);
  } else if (rows <= kBasisMaxDim && cols <= kBasisMaxDim) {
    // wide matrices orthogonalize the rows of M (fewer columns) when W and V fit in LDS together
    const bool accum = cols > rows && cols * rows + rows * rows <= kBasisLds;
    const int k = accum ? rows : cols;
    const int kk = accum ? cols : rows;  // rows of the working matrix: 16 * RPL >= kk
    launch_basis(accum, rpl_for(kk), threads_for_pairs(k), m->stream, m->work, rows, cols, m->work + slot, m->sv,
                 m->info);
  } else {
    hipLaunchKernelGGL(k_mps_svd, dim3(1), dim3(kSvdThreads), 0, m->stream, m->work, rows, cols, m->work + 3 * slot,
                       m->work + 4 * slot, m->work + slot, m->sv, m->work + 2 * slot, m->info);
  }
  MPS_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_mps_project, dim3(blocks_for(static_cast<uint64_t>(kmin) * cols)), dim3(kThreads), 0,
                     m->stream, m->work + slot, m->work, m->work + 2 * slot, rows, cols, kmin);
  MPS_HIP(hipGetLastError());
  int h_info = 0;
  m->h_sv.resize(kmin);
  MPS_HIP(hipMemcpyAsync(m->h_sv.data(), m->sv, sizeof(float) * kmin, hipMemcpyDeviceToHost, m->stream));
  MPS_HIP(hipMemcpyAsync(&h_info, m->info, sizeof(int), hipMemcpyDeviceToHost, m->stream));
  MPS_HIP(hipStreamSynchronize(m->stream));
  if (h_info < 0) return set_error(MSV_E_INVALID, "MPS SVD did not converge");
  for (int k = 0; k < kmin; ++k)
    if (!std::isfinite(m->h_sv[k])) return set_error(MSV_E_INVALID, "MPS SVD produced non-finite singular values");
  return MSV_OK;
}

// M = Q R for a center move: k_mps_qr_lds when M and Q fit in LDS (no host
// round trip: nothing is truncated), else the SVD factorization (R = S Vh).
int factor_move(msv_mps *m, int rows, int cols) {
  if (rows > kBasisMaxDim || cols > kBasisMaxDim ||
      static_cast<int64_t>(rows) * cols + static_cast<int64_t>(rows) * std::min(rows, cols) > kBasisLds)
    return factor(m, rows, cols);
  const size_t slot = static_cast<size_t>(4) * m->dmax * m->dmax;
  const int kmin = std::min(rows, cols);
  const int threads = std::min(kSvdThreads, std::max(64, (cols * kPairLanes + 63) / 64 * 64));
  C *q = m->work + slot, *r = m->work + 2 * slot;
  if (rows <= 16)
    hipLaunchKernelGGL(k_mps_qr_lds<1>, dim3(1), dim3(threads), 0, m->stream, m->work, rows, cols, q, r);
  else if (rows <= 32)
    hipLaunchKernelGGL(k_mps_qr_lds<2>, dim3(1), dim3(threads), 0, m->stream, m->work, rows, cols, q, r);
  else if (rows <= 64)
    hipLaunchKernelGGL(k_mps_qr_lds<4>, dim3(1), dim3(threads), 0, m->stream, m->work, rows, cols, q, r);
  else
    hipLaunchKernelGGL(k_mps_qr_lds<8>, dim3(1), dim3(threads), 0, m->stream, m->work, rows, cols, q, r);
  MPS_HIP(hipGetLastError());
  return MSV_OK;
}

// Moves the orthogonality center one site right (from i to i + 1) or left
// (from i to i - 1) without truncation: the reference's QR moves
// (mps_state.py:347-389), M = Q R with R carried into the neighbour.
int move_right(msv_mps *m, int i) {
  const int dl = m->dim[i], dr = m->dim[i + 1], dn = m->dim[i + 2];
  const int rows = 2 * dl, cols = dr, kmin = std::min(rows, cols);
  hipLaunchKernelGGL(k_mps_to_colmajor, dim3(blocks_for(static_cast<uint64_t>(rows) * cols)), dim3(kThreads), 0,
                     m->stream, site_ptr(m, i), m->work, rows, cols, 0);
  MPS_HIP(hipGetLastError());
  MPS_TRY(factor_move(m, rows, cols));
  const size_t slot = static_cast<size_t>(4) * m->dmax * m->dmax;
  hipLaunchKernelGGL(k_mps_absorb_right, dim3(blocks_for(static_cast<uint64_t>(rows) * kmin + kmin * 2 * dn)),
                     dim3(kThreads), 0, m->stream, m->work + slot, m->work + 2 * slot, site_ptr(m, i),
                     site_ptr(m, i + 1), m->spare, rows, cols, kmin, dn);
  MPS_HIP(hipGetLastError());
  std::swap(m->site[i + 1], m->spare);  // later writes to the old buffer are ordered after its reads
  m->dim[i + 1] = kmin;
  m->center = i + 1;
  return MSV_OK;
}

int move_left(msv_mps *m, int i) {
  const int dl = m->dim[i], dr = m->dim[i + 1], dp = m->dim[i - 1];
  // M = site^H, (2 dr) x dl
  const int rows = 2 * dr, cols = dl, kmin = std::min(rows, cols);
  hipLaunchKernelGGL(k_mps_to_colmajor, dim3(blocks_for(static_cast<uint64_t>(rows) * cols)), dim3(kThreads), 0,
                     m->stream, site_ptr(m, i), m->work, dl, 2 * dr, 1);
  MPS_HIP(hipGetLastError());
  MPS_TRY(factor_move(m, rows, cols));
  const size_t slot = static_cast<size_t>(4) * m->dmax * m->dmax;
  hipLaunchKernelGGL(k_mps_absorb_left, dim3(blocks_for(static_cast<uint64_t>(kmin) * rows + dp * 2 * kmin)),
                     dim3(kThreads), 0, m->stream, m->work + slot, m->work + 2 * slot, site_ptr(m, i),
                     site_ptr(m, i - 1), m->spare, dl, rows, kmin, dp);
  MPS_HIP(hipGetLastError());
  std::swap(m->site[i - 1], m->spare);
  m->dim[i] = kmin;
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
  MPS_TRY(factor(m, rows, cols, true));
  m->st.svd_calls += 1;
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
                     dim3(kThreads), 0, m->stream, m->work + slot, m->work + 2 * slot, site_ptr(m, i),
                     site_ptr(m, i + 1), dl, dr, kmin, r, scale);
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
  // dmax <= 1024: the global-memory Jacobi keeps a norm and a rank per column
  // of the two-site matrix (2 dmax columns at most) in LDS
  if (dmax < 1 || dmax > 1024) return set_error(MSV_E_INVALID, "MPS dmax must be in [1, 1024]");
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
      hipMalloc(reinterpret_cast<void **>(&m->spare_block), sizeof(C) * m->cap) != hipSuccess ||
      hipMalloc(reinterpret_cast<void **>(&m->work), sizeof(C) * slot * 5) != hipSuccess ||
      hipMalloc(reinterpret_cast<void **>(&m->sv), sizeof(float) * 4 * dmax) != hipSuccess ||
      hipMalloc(reinterpret_cast<void **>(&m->info), sizeof(int)) != hipSuccess ||
      hipMalloc(reinterpret_cast<void **>(&m->hh_beta), sizeof(double) * 2 * dmax) != hipSuccess ||
      hipMalloc(reinterpret_cast<void **>(&m->hh_piv), sizeof(int) * 2 * dmax) != hipSuccess)
    return cleanup(set_error(MSV_E_OOM, "MPS device allocation failed"));
  m->site.resize(n_qubits);
  for (int i = 0; i < n_qubits; ++i) m->site[i] = m->sites + static_cast<size_t>(i) * m->cap;
  m->spare = m->spare_block;
  const int rc = msv_mps_reset(m);
  if (rc != MSV_OK) return cleanup(rc);
  *out = m;
  return MSV_OK;
}

int msv_mps_destroy(msv_mps *m) {
  if (!m) return MSV_OK;
  // everything is released whatever fails; the first error is reported
  hipError_t e = hipSetDevice(m->device);
  if (m->stream) {
    const hipError_t es = hipStreamSynchronize(m->stream);
    if (e == hipSuccess) e = es;
  }
  for (void *p : {static_cast<void *>(m->sites), static_cast<void *>(m->spare_block), static_cast<void *>(m->work),
                  static_cast<void *>(m->sv), static_cast<void *>(m->info), static_cast<void *>(m->hh_beta),
                  static_cast<void *>(m->hh_piv)}) {
    const hipError_t ef = hipFree(p);
    if (e == hipSuccess) e = ef;
  }
  if (m->stream) {
    const hipError_t ed = hipStreamDestroy(m->stream);
    if (e == hipSuccess) e = ed;
  }
  delete m;
  return e == hipSuccess ? MSV_OK : fail_hip(e, "MPS destroy");
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
  // two environments (dn x dn each) and the intermediate T (dm x 2 x dn)
  MPS_HIP(hipMallocAsync(reinterpret_cast<void **>(&env), sizeof(double2) * 4 * emax, m->stream));
  double2 *tmid = env + 2 * emax;
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
    hipLaunchKernelGGL(k_mps_transfer_a, dim3(blocks_for(static_cast<uint64_t>(dm) * 2 * dn)), dim3(kThreads), 0,
                       m->stream, env + cur * emax, site_ptr(m, i), tmid, dm, dn);
    hipLaunchKernelGGL(k_mps_transfer_b, dim3(blocks_for(static_cast<uint64_t>(dn) * dn)), dim3(kThreads), 0,
                       m->stream, tmid, site_ptr(m, i), env + (1 - cur) * emax, dm, dn, O);
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
