/* CPU restatement (FP64, OpenMP) of the reference time loop, for the CPU
 * baseline timed beside the GPU.  Test infrastructure: only tests/ and
 * bench.py's CPU-baseline leg load it; the HIP library never does.
 *
 * It follows core_shell.jl op for op, as oracle/core_shell_oracle.py does:
 *   make_grid         core_shell.jl:17-25   (k, derivative k with the Nyquist
 *                                             mode zeroed, k2 with it kept)
 *   geometry          :39-43   h = phi^3 (10 - 15 phi + 6 phi^2), g, M
 *   bulk_energy       :45-49
 *   bulk_derivative   :51-55
 *   chemical_potential:57-60   2 transforms
 *   transport_rate    :62-70   7 transforms (mu_hat, grad_x, grad_y, the two
 *                                             flux transforms)
 *   step              :72-76   transport_rate + 2 transforms
 *   energy            :78-82   1 transform
 *   run_simulation    :84-103  per iteration transport_rate, history row,
 *                              finiteness check, step: 17 complex N x N 2D
 *                              transforms, like the reference.
 * The transforms follow FFTW's convention (unnormalised forward e^{-i}, 1/N^2
 * inverse), here as radix-2 Cooley-Tukey with twiddles computed in long
 * double, so N must be a power of two.  Arrays are column-major like the
 * Julia matrices: a[i + n j], kx varying along i.  Sums run in OpenMP
 * reductions, so the last bits depend on the thread count.
 */
#include <complex.h>
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

typedef double complex cd;

enum { CS_OK = 0, CS_ERR_ARG = 1, CS_ERR_MEM = 2, CS_ERR_NONFINITE = 3 };

static const double PI = 3.141592653589793;   /* the double nearest pi, as numpy's */

typedef struct {
  int n;
  cd* tw;    /* per stage of length len = 2, 4, .., n: e^{-2 pi i k / len}, k < len/2, at tw + len/2 - 1 */
  int* rev;  /* bit reversal of 0..n-1 */
} plan_t;

static int plan_init(plan_t* p, int n) {
  p->n = n;
  p->tw = malloc(sizeof(cd) * (size_t)(n > 1 ? n - 1 : 1));
  p->rev = malloc(sizeof(int) * (size_t)n);
  if (!p->tw || !p->rev) return CS_ERR_MEM;
  const long double two_pi = 6.283185307179586476925286766559L;
  for (int len = 2; len <= n; len <<= 1)
    for (int k = 0; k < len / 2; ++k) {
      /* the same value as W_n^(k n/len): the exponent reduced to k/len exactly */
      const long double a = -two_pi * (long double)k / (long double)len;
      p->tw[len / 2 - 1 + k] = (double)cosl(a) + I * (double)sinl(a);
    }
  int bits = 0;
  while ((1 << bits) < n) ++bits;
  for (int i = 0; i < n; ++i) {
    int r = 0;
    for (int b = 0; b < bits; ++b) r |= ((i >> b) & 1) << (bits - 1 - b);
    p->rev[i] = r;
  }
  return CS_OK;
}

static void plan_free(plan_t* p) {
  free(p->tw);
  free(p->rev);
}

/* In-place length-n transform of x (unnormalised; inv flips the sign). */
static void fft1(const plan_t* p, cd* x, int inv) {
  const int n = p->n;
  for (int i = 0; i < n; ++i) {
    const int r = p->rev[i];
    if (r > i) {
      const cd t = x[i];
      x[i] = x[r];
      x[r] = t;
    }
  }
  for (int len = 2; len <= n; len <<= 1) {
    const int half = len >> 1;
    const cd* tw = p->tw + half - 1;
    const double sgn = inv ? -1.0 : 1.0;
    for (int i = 0; i < n; i += len) {
      for (int k = 0; k < half; ++k) {
        const double wr = creal(tw[k]), wi = sgn * cimag(tw[k]);
        const double xr = creal(x[i + k + half]), xi = cimag(x[i + k + half]);
        const cd v = (xr * wr - xi * wi) + I * (xr * wi + xi * wr);
        const cd u = x[i + k];
        x[i + k] = u + v;
        x[i + k + half] = u - v;
      }
    }
  }
}

/* 2D transform of a (n x n, column-major) in place: along i (contiguous
 * columns), then along j in tiles of TB rows gathered into a per-thread
 * buffer; the inverse is scaled by 1/n^2. */
#define TB 16
static void fft2(const plan_t* p, cd* a, int inv) {
  const int n = p->n;
#pragma omp parallel for schedule(static)
  for (int j = 0; j < n; ++j) fft1(p, a + (size_t)n * j, inv);
#pragma omp parallel
  {
    cd* buf = malloc(sizeof(cd) * (size_t)n * TB);
#pragma omp for schedule(static)
    for (int i0 = 0; i0 < n; i0 += TB) {
      const int tb = n - i0 < TB ? n - i0 : TB;
      for (int j = 0; j < n; ++j)
        for (int t = 0; t < tb; ++t) buf[(size_t)t * n + j] = a[(size_t)n * j + i0 + t];
      for (int t = 0; t < tb; ++t) fft1(p, buf + (size_t)t * n, inv);
      for (int j = 0; j < n; ++j)
        for (int t = 0; t < tb; ++t) a[(size_t)n * j + i0 + t] = buf[(size_t)t * n + j];
    }
    free(buf);
  }
  if (inv) {
    const double s = 1.0 / ((double)n * (double)n);
    const size_t nn = (size_t)n * n;
#pragma omp parallel for schedule(static)
    for (size_t q = 0; q < nn; ++q) a[q] *= s;
  }
}

typedef struct {
  double A, B, chi, P, ca, cb, cc, kappa, mobility, Q;
} params_t;

typedef struct {
  int n;
  double spacing;
  params_t p;
  plan_t plan;
  double *k, *dk;        /* wavenumbers, derivative wavenumbers (n each) */
  double *h, *g, *M;     /* geometry (n^2 each) */
  cd *w0, *w1, *w2;      /* transform work arrays (n^2 each) */
  double *r0, *r1;       /* real work arrays (n^2 each) */
} sim_t;

#define K2(s, i, j) ((s)->k[i] * (s)->k[i] + (s)->k[j] * (s)->k[j])

static void sim_free(sim_t* s) {
  plan_free(&s->plan);
  free(s->k); free(s->dk); free(s->h); free(s->g); free(s->M);
  free(s->w0); free(s->w1); free(s->w2); free(s->r0); free(s->r1);
}

static int sim_init(sim_t* s, int n, double spacing, const double* prm, const double* phi) {
  memset(s, 0, sizeof(*s));
  s->n = n;
  s->spacing = spacing;
  memcpy(&s->p, prm, sizeof(params_t));
  if (plan_init(&s->plan, n)) return CS_ERR_MEM;
  const size_t nn = (size_t)n * n;
  s->k = malloc(sizeof(double) * n);
  s->dk = malloc(sizeof(double) * n);
  s->h = malloc(sizeof(double) * nn);
  s->g = malloc(sizeof(double) * nn);
  s->M = malloc(sizeof(double) * nn);
  s->w0 = malloc(sizeof(cd) * nn);
  s->w1 = malloc(sizeof(cd) * nn);
  s->w2 = malloc(sizeof(cd) * nn);
  s->r0 = malloc(sizeof(double) * nn);
  s->r1 = malloc(sizeof(double) * nn);
  if (!s->k || !s->dk || !s->h || !s->g || !s->M || !s->w0 || !s->w1 || !s->w2 || !s->r0 || !s->r1)
    return CS_ERR_MEM;
  /* make_grid: k = 2 pi fftfreq(n, 1/spacing) (core_shell.jl:18-23) */
  const double f = (1.0 / spacing) / n;
  for (int m = 0; m < n; ++m) {
    const double ms = (m < (n + 1) / 2) ? (double)m : (double)(m - n);
    s->k[m] = (2.0 * PI) * (ms * f);
    s->dk[m] = s->k[m];
  }
  if (n % 2 == 0) s->dk[n / 2] = 0.0;
  /* geometry (core_shell.jl:39-43) */
#pragma omp parallel for schedule(static)
  for (size_t q = 0; q < nn; ++q) {
    const double x = phi[q];
    s->h[q] = pow(x, 3) * (10 - 15 * x + 6 * pow(x, 2));
    s->g[q] = pow(x, 2) * pow(1 - x, 2);
    s->M[q] = s->p.mobility * s->h[q];
  }
  return CS_OK;
}

static double bulk_derivative(const sim_t* s, double c, size_t q) {
  const params_t* p = &s->p;
  const double h = s->h[q], g = s->g[q];
  return 2 * p->A * (1 - h) * (c - p->ca) + 2 * p->B * h * (c - p->cb) * (c - p->cc) * (2 * c - p->cb - p->cc) -
         p->chi * p->P * g;
}

static double bulk_energy(const sim_t* s, double c, size_t q) {
  const params_t* p = &s->p;
  const double h = s->h[q], g = s->g[q];
  return p->A * (1 - h) * pow(c - p->ca, 2) + p->B * h * pow(c - p->cb, 2) * pow(c - p->cc, 2) +
         p->P * (1 - p->chi * c) * g;
}

/* mu = f'(c) + Re ifft(2 kappa k2 fft(c)) into r0 (core_shell.jl:57-60) */
static void chemical_potential(sim_t* s, const double* c) {
  const int n = s->n;
  const size_t nn = (size_t)n * n;
#pragma omp parallel for schedule(static)
  for (size_t q = 0; q < nn; ++q) s->w0[q] = c[q];
  fft2(&s->plan, s->w0, 0);
  const double tk = 2 * s->p.kappa;
#pragma omp parallel for schedule(static)
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) s->w0[(size_t)n * j + i] *= tk * K2(s, i, j);
  fft2(&s->plan, s->w0, 1);
#pragma omp parallel for schedule(static)
  for (size_t q = 0; q < nn; ++q) s->r0[q] = bulk_derivative(s, c[q], q) + creal(s->w0[q]);
}

/* rate (k-space) into w2 and the dissipation (core_shell.jl:62-70) */
static double transport_rate(sim_t* s, const double* c) {
  const int n = s->n;
  const size_t nn = (size_t)n * n;
  chemical_potential(s, c);
  /* mu_hat in w1 */
#pragma omp parallel for schedule(static)
  for (size_t q = 0; q < nn; ++q) s->w1[q] = s->r0[q];
  fft2(&s->plan, s->w1, 0);
  /* grad_x = Re ifft(i kx mu_hat) in r0, grad_y in r1 */
#pragma omp parallel for schedule(static)
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) s->w0[(size_t)n * j + i] = I * s->dk[i] * s->w1[(size_t)n * j + i];
  fft2(&s->plan, s->w0, 1);
#pragma omp parallel for schedule(static)
  for (size_t q = 0; q < nn; ++q) s->r0[q] = creal(s->w0[q]);
#pragma omp parallel for schedule(static)
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) s->w0[(size_t)n * j + i] = I * s->dk[j] * s->w1[(size_t)n * j + i];
  fft2(&s->plan, s->w0, 1);
#pragma omp parallel for schedule(static)
  for (size_t q = 0; q < nn; ++q) s->r1[q] = creal(s->w0[q]);
  /* divergence_hat = i kx fft(M gx) + i ky fft(M gy) */
  double d = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : d)
  for (size_t q = 0; q < nn; ++q) {
    d += s->M[q] * (s->r0[q] * s->r0[q] + s->r1[q] * s->r1[q]);
    s->w0[q] = s->M[q] * s->r0[q];
    s->w1[q] = s->M[q] * s->r1[q];
  }
  fft2(&s->plan, s->w0, 0);
  fft2(&s->plan, s->w1, 0);
#pragma omp parallel for schedule(static)
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      const size_t q = (size_t)n * j + i;
      s->w2[q] = I * s->dk[i] * s->w0[q] + I * s->dk[j] * s->w1[q];
    }
  return d * s->spacing * s->spacing;
}

/* F (core_shell.jl:78-82) */
static double energy(sim_t* s, const double* c) {
  const int n = s->n;
  const size_t nn = (size_t)n * n;
#pragma omp parallel for schedule(static)
  for (size_t q = 0; q < nn; ++q) s->w0[q] = c[q];
  fft2(&s->plan, s->w0, 0);
  double ge = 0.0, fb = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : ge, fb)
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      const size_t q = (size_t)n * j + i;
      const double a = cabs(s->w0[q]);
      ge += K2(s, i, j) * (a * a);
      fb += bulk_energy(s, c[q], q);
    }
  const double grad = s->p.kappa * ge / (double)nn;
  return (fb + grad) * s->spacing * s->spacing;
}

/* c <- Re ifft(fft(c) + dt rate / (1 + 2 Q kappa k2^2 dt)) (core_shell.jl:72-76) */
static void step(sim_t* s, double* c, double dt) {
  const int n = s->n;
  const size_t nn = (size_t)n * n;
  (void)transport_rate(s, c);
#pragma omp parallel for schedule(static)
  for (size_t q = 0; q < nn; ++q) s->w0[q] = c[q];
  fft2(&s->plan, s->w0, 0);
  const double qk = 2 * s->p.Q * s->p.kappa;
#pragma omp parallel for schedule(static)
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      const size_t q = (size_t)n * j + i;
      const double k2 = K2(s, i, j);
      s->w0[q] = s->w0[q] + dt * s->w2[q] / (1 + qk * (k2 * k2) * dt);
    }
  fft2(&s->plan, s->w0, 1);
#pragma omp parallel for schedule(static)
  for (size_t q = 0; q < nn; ++q) c[q] = creal(s->w0[q]);
}

/* run_simulation (core_shell.jl:84-103) on `threads` OpenMP threads: c_out
 * gets c(steps), hist (steps+1 rows of 6, row-major) the history rows
 * [t, mean c, F, min c, max c, dissipation] of c_s before it is advanced.
 * Returns CS_ERR_NONFINITE with *bad_step set at the first non-finite c_s. */
int cs_run(int n, double spacing, const double* prm, double dt, long steps, const double* c0,
           const double* phi, double* c_out, double* hist, int threads, long* bad_step) {
  if (n < 2 || (n & (n - 1)) || steps < 0 || !prm || !c0 || !phi || !c_out || !hist || threads < 1)
    return CS_ERR_ARG;
  omp_set_num_threads(threads);
  sim_t s;
  int rc = sim_init(&s, n, spacing, prm, phi);
  if (rc) {
    sim_free(&s);
    return rc;
  }
  const size_t nn = (size_t)n * n;
  memcpy(c_out, c0, sizeof(double) * nn);
  if (bad_step) *bad_step = -1;
  for (long st = 0; st <= steps; ++st) {
    const double d = transport_rate(&s, c_out);
    double sum = 0.0, mn = INFINITY, mx = -INFINITY;
    int bad = 0;
#pragma omp parallel for schedule(static) reduction(+ : sum, bad) reduction(min : mn) reduction(max : mx)
    for (size_t q = 0; q < nn; ++q) {
      const double v = c_out[q];
      sum += v;
      mn = fmin(mn, v);
      mx = fmax(mx, v);
      bad += isfinite(v) ? 0 : 1;
    }
    double* row = hist + 6 * st;
    row[0] = st * dt;
    row[1] = sum / (double)nn;
    row[2] = energy(&s, c_out);
    row[3] = mn;
    row[4] = mx;
    row[5] = d;
    if (bad) {
      if (bad_step) *bad_step = st;
      sim_free(&s);
      return CS_ERR_NONFINITE;
    }
    if (st < steps) step(&s, c_out, dt);
  }
  sim_free(&s);
  return CS_OK;
}
