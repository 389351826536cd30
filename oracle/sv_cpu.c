/*
 * sv_cpu.c -- CPU restatement of the reference statevector path
 * (TEST INFRASTRUCTURE: the checker and the timed CPU baseline, never the
 * product).
 *
 * On a CPU device the reference disables every Metal fusion
 * (execution.py:167-176) and runs gate by gate through
 * StateVectorSimulator (paths relative to the reference's src/mettleq):
 *   cpu_apply_1q      sim.py:75-87    apply_single: (A,2,B) slices, complex64
 *   cpu_apply_diag1   sim.py:94-117   apply_diagonal, k = 1
 *   cpu_apply_diag2   sim.py:118-129  apply_diagonal, k = 2 (CZ/CP/CRZ/ZZPHASE)
 *   cpu_apply_cu      sim.py:140-173  apply_controlled_single (CNOT, CRX, ...)
 *   cpu_apply_swap    sim.py:215-224  apply_swap_wires
 *   cpu_qft_ladder    sim.py:485-510  _fused_phase_ladder (one diagonal per QFT stage)
 * Arithmetic is complex64 like the MLX arrays it restates. Loops run under
 * OpenMP; every amplitude is read and written once per gate.
 *
 * The sampling functions restate this engine's own algorithm (Philox4x32-10
 * + fixed-order FP64 block CDF) so that device samples can be checked bit
 * for bit; the reference samples with mx.random.categorical (sim.py:405-413).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
  float re, im;
} c64;

static inline c64 cmulf(c64 a, float br, float bi) {
  c64 r;
  r.re = a.re * br - a.im * bi;
  r.im = a.re * bi + a.im * br;
  return r;
}

int cpu_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void cpu_set_threads(int t) {
#ifdef _OPENMP
  if (t > 0) omp_set_num_threads(t);
#else
  (void)t;
#endif
}

/* Qubit q is index bit n-1-q. */
void cpu_apply_1q(c64 *st, int n, int q, const double *m) {
  const int64_t stride = (int64_t)1 << (n - 1 - q);
  const int64_t npairs = (int64_t)1 << (n - 1);
  const float m0 = (float)m[0], m1 = (float)m[1], m2 = (float)m[2], m3 = (float)m[3];
  const float m4 = (float)m[4], m5 = (float)m[5], m6 = (float)m[6], m7 = (float)m[7];
#pragma omp parallel for schedule(static)
  for (int64_t p = 0; p < npairs; ++p) {
    const int64_t i0 = ((p / stride) * 2 * stride) + (p % stride);
    const int64_t i1 = i0 + stride;
    const c64 a = st[i0], b = st[i1];
    c64 x, y;
    x.re = (m0 * a.re - m1 * a.im) + (m2 * b.re - m3 * b.im);
    x.im = (m0 * a.im + m1 * a.re) + (m2 * b.im + m3 * b.re);
    y.re = (m4 * a.re - m5 * a.im) + (m6 * b.re - m7 * b.im);
    y.im = (m4 * a.im + m5 * a.re) + (m6 * b.im + m7 * b.re);
    st[i0] = x;
    st[i1] = y;
  }
}

void cpu_apply_diag1(c64 *st, int n, int q, const double *d) {
  const int64_t N = (int64_t)1 << n;
  const int sh = n - 1 - q;
  const float d0r = (float)d[0], d0i = (float)d[1], d1r = (float)d[2], d1i = (float)d[3];
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < N; ++i) {
    if ((i >> sh) & 1) st[i] = cmulf(st[i], d1r, d1i);
    else st[i] = cmulf(st[i], d0r, d0i);
  }
}

/* Diagonal over (a, b) with a as the most significant bit of the 4-entry
 * table (sim.py:118-129 reorders when a > b; indexing by bit values here is
 * the same thing). */
void cpu_apply_diag2(c64 *st, int n, int a, int b, const double *d) {
  const int64_t N = (int64_t)1 << n;
  const int sa = n - 1 - a, sb = n - 1 - b;
  float dr[4], di[4];
  for (int k = 0; k < 4; ++k) {
    dr[k] = (float)d[2 * k];
    di[k] = (float)d[2 * k + 1];
  }
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < N; ++i) {
    const int sel = (int)(((i >> sa) & 1) * 2 + ((i >> sb) & 1));
    st[i] = cmulf(st[i], dr[sel], di[sel]);
  }
}

void cpu_apply_cu(c64 *st, int n, int c, int t, const double *m) {
  const int64_t N = (int64_t)1 << n;
  const int sc = n - 1 - c, stt = n - 1 - t;
  const int64_t tb = (int64_t)1 << stt;
  const float m0 = (float)m[0], m1 = (float)m[1], m2 = (float)m[2], m3 = (float)m[3];
  const float m4 = (float)m[4], m5 = (float)m[5], m6 = (float)m[6], m7 = (float)m[7];
#pragma omp parallel for schedule(static)
  for (int64_t i0 = 0; i0 < N; ++i0) {
    if (((i0 >> sc) & 1) == 0 || ((i0 >> stt) & 1) != 0) continue;
    const int64_t i1 = i0 | tb;
    const c64 a = st[i0], b = st[i1];
    c64 x, y;
    x.re = (m0 * a.re - m1 * a.im) + (m2 * b.re - m3 * b.im);
    x.im = (m0 * a.im + m1 * a.re) + (m2 * b.im + m3 * b.re);
    y.re = (m4 * a.re - m5 * a.im) + (m6 * b.re - m7 * b.im);
    y.im = (m4 * a.im + m5 * a.re) + (m6 * b.im + m7 * b.re);
    st[i0] = x;
    st[i1] = y;
  }
}

void cpu_apply_swap(c64 *st, int n, int a, int b) {
  if (a == b) return;
  const int64_t N = (int64_t)1 << n;
  const int sa = n - 1 - a, sb = n - 1 - b;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < N; ++i) {
    const int ba = (int)((i >> sa) & 1), bb = (int)((i >> sb) & 1);
    if (ba == 1 && bb == 0) {
      const int64_t j = (i & ~((int64_t)1 << sa)) | ((int64_t)1 << sb);
      const c64 t = st[i];
      st[i] = st[j];
      st[j] = t;
    }
  }
}

/* One QFT stage's controlled-phase ladder as a single diagonal:
 * phase = bit_j * sum_{k>j} bit_k * sign*pi/2^(k-j) over the ordered wire
 * list, each factor rounded to complex64 and multiplied in like the
 * broadcast product of sim.py:485-510. */
void cpu_qft_ladder(c64 *st, int n, const int *wires, int nw, int j, double sign) {
  const int64_t N = (int64_t)1 << n;
  float fr[64], fi[64];
  for (int k = j + 1; k < nw; ++k) {
    const double phi = sign * M_PI / ldexp(1.0, k - j);
    fr[k] = (float)cos(phi);
    fi[k] = (float)sin(phi);
  }
  const int sj = n - 1 - wires[j];
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < N; ++i) {
    if (((i >> sj) & 1) == 0) continue;
    float pr = 1.0f, pi = 0.0f;
    for (int k = j + 1; k < nw; ++k) {
      if ((i >> (n - 1 - wires[k])) & 1) {
        const float nr = pr * fr[k] - pi * fi[k];
        pi = pr * fi[k] + pi * fr[k];
        pr = nr;
      }
    }
    st[i] = cmulf(st[i], pr, pi);
  }
}

/* ---------------- sampling restatement ---------------- */

static void philox(uint32_t c[4], uint32_t k0, uint32_t k1) {
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
    const uint32_t n1 = (uint32_t)p1;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
    const uint32_t n3 = (uint32_t)p0;
    c[0] = n0;
    c[1] = n1;
    c[2] = n2;
    c[3] = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

void cpu_philox4x32_10(const uint32_t *ctr, const uint32_t *key, uint32_t *out) {
  uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
  philox(c, key[0], key[1]);
  memcpy(out, c, sizeof(c));
}

double cpu_philox_u01(uint64_t seed, uint64_t stream_id, uint64_t counter) {
  uint32_t c[4] = {(uint32_t)counter, (uint32_t)(counter >> 32), (uint32_t)stream_id,
                   (uint32_t)(stream_id >> 32)};
  philox(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  const uint64_t a = c[0] >> 5, b = c[1] >> 6;
  return ((double)a * 67108864.0 + (double)b) * (1.0 / 9007199254740992.0);
}

#define BLK 4096
#define LN 64

static inline double prob_at(const c64 *st, int64_t N, int64_t i) {
  if (i >= N) return 0.0;
  const double re = (double)st[i].re, im = (double)st[i].im;
  const double a = re * re;
  const double b = im * im;
  return a + b;
}

static void lane_sums(const c64 *st, int64_t N, int64_t b, double *ls) {
  for (int l = 0; l < LN; ++l) {
    double s = 0.0;
    for (int j = 0; j < BLK / LN; ++j) s = s + prob_at(st, N, b * BLK + (int64_t)j * LN + l);
    ls[l] = s;
  }
}

void cpu_block_sums(const c64 *st, int n, double *out) {
  const int64_t N = (int64_t)1 << n;
  const int64_t nb = (N + BLK - 1) / BLK;
#pragma omp parallel for schedule(static)
  for (int64_t b = 0; b < nb; ++b) {
    double ls[LN];
    lane_sums(st, N, b, ls);
    double t = 0.0;
    for (int l = 0; l < LN; ++l) t = t + ls[l];
    out[b] = t;
  }
}

/* Returns 0 on success, 1 on a zero-norm state. */
int cpu_sample(const c64 *st, int n, uint64_t shots, uint64_t seed, uint64_t stream_id,
               uint64_t *out_index) {
  const int64_t N = (int64_t)1 << n;
  const int64_t nb = (N + BLK - 1) / BLK;
  double *bs = (double *)malloc(sizeof(double) * nb);
  double *prefix = (double *)malloc(sizeof(double) * (nb + 1));
  cpu_block_sums(st, n, bs);
  prefix[0] = 0.0;
  for (int64_t b = 0; b < nb; ++b) prefix[b + 1] = prefix[b] + bs[b];
  const double total = prefix[nb];
  if (!(total > 0.0)) {
    free(bs);
    free(prefix);
    return 1;
  }
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < (int64_t)shots; ++k) {
    double t = cpu_philox_u01(seed, stream_id, (uint64_t)k) * total;
    if (t >= total) t = nextafter(total, 0.0);
    /* first b with prefix[b+1] > t */
    int64_t lo = 0, hi = nb;
    while (lo < hi) {
      const int64_t mid = (lo + hi) / 2;
      if (prefix[mid + 1] > t) hi = mid;
      else lo = mid + 1;
    }
    int64_t b = lo < nb ? lo : nb - 1;
    const double r = t - prefix[b];
    double ls[LN];
    lane_sums(st, N, b, ls);
    double c = 0.0, r2 = 0.0;
    int L = -1, last = -1;
    for (int l = 0; l < LN; ++l) {
      const double v = ls[l];
      if (v > 0.0) last = l;
      const double cn = c + v;
      if (L < 0 && cn > r) {
        L = l;
        r2 = r - c;
      }
      c = cn;
    }
    if (L < 0) {
      L = last < 0 ? 0 : last;
      r2 = INFINITY;
    }
    double c2 = 0.0;
    int J = -1, lastj = -1;
    for (int j = 0; j < BLK / LN; ++j) {
      const double v = prob_at(st, N, b * BLK + (int64_t)j * LN + L);
      if (v > 0.0) lastj = j;
      const double cn = c2 + v;
      if (J < 0 && cn > r2) J = j;
      c2 = cn;
    }
    if (J < 0) J = lastj < 0 ? 0 : lastj;
    out_index[k] = (uint64_t)(b * BLK + (int64_t)J * LN + L);
  }
  free(bs);
  free(prefix);
  return 0;
}
