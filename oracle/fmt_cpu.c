/*
 * FP64 CPU restatement of the new-API White-Bear minimization -- TEST
 * INFRASTRUCTURE: the CPU baseline that bench.py times and tests/ check; the
 * product path (deft_amd) never runs it.
 *
 * It follows the same reference lines as oracle/fmt_oracle.py, in plain C on
 * one thread, with the new API's dataflow (cached energies; one r2c and five
 * c2r transforms per energy, six and six per energy + gradient):
 *   - layout, transforms and k-vectors      src/Grid.h:37-42, src/Grid.cpp:470-478,
 *                                           src/ReciprocalOperators.h:20-41
 *   - weights with smearing and series      src/ReciprocalOperators.h:90-235, src/haskell/FMT.hs:21-65
 *   - White-Bear Phi and partials           src/HardSpheres.cpp:137-159, src/haskell/WhiteBear.hs:47-60,114-121
 *   - ideal gas, Vext, mu, Veff chain rule  src/IdealGas.cpp:54-72, src/EffectivePotentialToDensity.cpp:28-90
 *   - new-API preconditioner grad/n         src/haskell/NewCode.hs:229-242
 *   - Minimize (PR-CG, quadratic line search, error estimate)  src/new/Minimize.cpp:9-313
 * Transforms: the project's own complex Stockham FFT (radices 4, 2, 3, 5); the
 * z pass transforms two real lines per complex FFT.  Weight tables are built
 * once per grid.
 *
 * usage:
 *   fmt_cpu eval <in> <out>
 *       in: Nx Ny Nz, cell[9], R kT mu, then Vext[N] and Veff[N] (text or binary
 *       as below); out: energy and the six terms, grad[N], pgrad[N] (binary doubles).
 *       The input file is "FMTC" + int32 Nx Ny Nz + 12 doubles + 2N doubles.
 *   fmt_cpu solute <N> <L> <eta> <radius> <precision> <maxiter> <sample_seconds>
 *       the 128^3 solute workload's shape: cubic cell L, hard solute of excluded-centre
 *       radius `radius` at the origin, mu from the bulk, Veff = -kT ln n_b, preconditioned
 *       Minimize to absolute precision `precision` (0: relative 1e-12); stops after
 *       sample_seconds (0: no limit).  Prints one JSON line.
 */
#define _POSIX_C_SOURCE 200809L
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

typedef struct {
  double re, im;
} cplx;

static double now(void) {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return t.tv_sec + 1e-9 * t.tv_nsec;
}

/* ---------------------------------------------------------------- FFT --- */
typedef struct {
  int n, nstage, radix[32];
  cplx *tw; /* exp(-2 pi i m / n), m < n */
  cplx *buf;
} fft_plan;

static int fft_init(fft_plan *p, int n) {
  p->n = n;
  p->nstage = 0;
  int m = n;
  while (m > 1) {
    int r = (m % 4 == 0) ? 4 : (m % 2 == 0) ? 2 : (m % 3 == 0) ? 3 : (m % 5 == 0) ? 5 : 0;
    if (!r || p->nstage == 32) return 0;
    p->radix[p->nstage++] = r;
    m /= r;
  }
  p->tw = malloc(sizeof(cplx) * n);
  p->buf = malloc(sizeof(cplx) * n);
  for (int k = 0; k < n; ++k) {
    const double a = 2 * M_PI * (double)k / (double)n;
    p->tw[k].re = cos(a);
    p->tw[k].im = -sin(a);
  }
  return 1;
}

/* In place, forward (exp(-2 pi i jk/n)) or inverse (exp(+...), unnormalised).
 * Stockham autosort, decimation in frequency: each stage does radix-r
 * butterflies over the sub-transforms of stride s and applies the twiddle
 * exp(-2 pi i j u / (r m)) to output u of butterfly j. */
static void fft_run(const fft_plan *p, cplx *x, int inverse) {
  const int n = p->n;
  cplx *a = x, *b = p->buf;
  int s = 1; /* stride of the current sub-transforms */
  int len = n;
  for (int st = 0; st < p->nstage; ++st) {
    const int r = p->radix[st];
    const int m = len / r; /* butterflies per group */
    for (int j = 0; j < m; ++j) {
      for (int k = 0; k < s; ++k) {
        cplx in[5];
        for (int q = 0; q < r; ++q) in[q] = a[k + s * (j + q * m)];
        /* twiddle exp(-+2 pi i j q / (r m)) applied after the r-point DFT */
        cplx outv[5];
        if (r == 2) {
          outv[0].re = in[0].re + in[1].re;
          outv[0].im = in[0].im + in[1].im;
          outv[1].re = in[0].re - in[1].re;
          outv[1].im = in[0].im - in[1].im;
        } else if (r == 4) {
          const cplx t0 = {in[0].re + in[2].re, in[0].im + in[2].im};
          const cplx t1 = {in[0].re - in[2].re, in[0].im - in[2].im};
          const cplx t2 = {in[1].re + in[3].re, in[1].im + in[3].im};
          cplx t3 = {in[1].re - in[3].re, in[1].im - in[3].im};
          /* t3 * (-i) forward, * (+i) inverse */
          const cplx t3r = inverse ? (cplx){-t3.im, t3.re} : (cplx){t3.im, -t3.re};
          outv[0].re = t0.re + t2.re;
          outv[0].im = t0.im + t2.im;
          outv[2].re = t0.re - t2.re;
          outv[2].im = t0.im - t2.im;
          outv[1].re = t1.re + t3r.re;
          outv[1].im = t1.im + t3r.im;
          outv[3].re = t1.re - t3r.re;
          outv[3].im = t1.im - t3r.im;
        } else {
          for (int u = 0; u < r; ++u) {
            double re = 0, im = 0;
            for (int q = 0; q < r; ++q) {
              const int e = (u * q) % r;
              const cplx w = p->tw[e * (n / r)];
              const double wi = inverse ? -w.im : w.im;
              re += in[q].re * w.re - in[q].im * wi;
              im += in[q].re * wi + in[q].im * w.re;
            }
            outv[u].re = re;
            outv[u].im = im;
          }
        }
        for (int u = 0; u < r; ++u) {
          const int e = j * u * s; /* exp(-2 pi i j u / (r m)) = tw[j u n / (r m)] */
          const cplx w = p->tw[(e % n)];
          const double wi = inverse ? -w.im : w.im;
          cplx o = outv[u];
          if (e) {
            const double re = o.re * w.re - o.im * wi;
            o.im = o.re * wi + o.im * w.re;
            o.re = re;
          }
          b[k + s * (u + r * j)] = o;
        }
      }
    }
    cplx *t = a;
    a = b;
    b = t;
    s *= r;
    len = m;
  }
  if (a != x) memcpy(x, a, sizeof(cplx) * n);
}

/* ------------------------------------------------------------- grids --- */
typedef struct {
  int Nx, Ny, Nz, Nkz;
  long long N, Nk;
  double cell[9], recip[9], volume, dV;
  double R, kT, mu, logNQ;
  fft_plan px, py, pz;
  double *w3, *w2, *wk; /* wk: 3 per element, the vector weight's i-coefficient times k_j */
  cplx *spec[6];
  cplx *line;
  double *vext;
  int nevE, nevG;
  double tE, tG;
} grid;

static void reciprocal(const double *c, double *b) {
  const double det = c[0] * (c[4] * c[8] - c[5] * c[7]) - c[1] * (c[3] * c[8] - c[5] * c[6]) +
                     c[2] * (c[3] * c[7] - c[4] * c[6]);
  double inv[9];
  inv[0] = (c[4] * c[8] - c[5] * c[7]) / det;
  inv[1] = (c[2] * c[7] - c[1] * c[8]) / det;
  inv[2] = (c[1] * c[5] - c[2] * c[4]) / det;
  inv[3] = (c[5] * c[6] - c[3] * c[8]) / det;
  inv[4] = (c[0] * c[8] - c[2] * c[6]) / det;
  inv[5] = (c[2] * c[3] - c[0] * c[5]) / det;
  inv[6] = (c[3] * c[7] - c[4] * c[6]) / det;
  inv[7] = (c[1] * c[6] - c[0] * c[7]) / det;
  inv[8] = (c[0] * c[4] - c[1] * c[3]) / det;
  for (int j = 0; j < 3; ++j)
    for (int k = 0; k < 3; ++k) b[j * 3 + k] = 2 * M_PI * inv[k * 3 + j];
}

static double nQ(double kT) { return pow(18.01528 * 1822.8885 * kT / (2 * M_PI), 1.5); }

static int wrapi(int m, int n) { return m > n / 2 ? m - n : m; }

/* w3, w2, wv of one |k| (the oracle's weights()) */
static void fm_weights(double k, double R, double dr, double *w3, double *w2, double *wv) {
  const double kR = k * R, kR2 = kR * kR;
  if (kR <= 1e-3) {
    *w3 = (4 * M_PI / 3) * R * R * R * (kR2 * kR2 * kR2 * (-1.0 / 15120) + kR2 * kR2 * (1.0 / 280) + kR2 * (-1.0 / 10) + 1);
    *w2 = 4 * M_PI * R * R * (-kR2 * kR2 * kR2 * (1.0 / 120 / 6 / 7) + kR2 * kR2 * (1.0 / 120) - kR2 * (1.0 / 6) + 1);
    *wv = 4 * M_PI * R * R * R * (kR2 * kR2 * (1.0 / 5040 - 1.0 / 720) + kR2 * (1.0 / 24 - 1.0 / 120) - 1.0 / 3);
    return;
  }
  const double s = exp(-6.0 * k * k * dr * dr);
  *w3 = s * 4 * M_PI * (sin(kR) - kR * cos(kR)) / (k * k * k);
  *w2 = s * 4 * M_PI * R * sin(kR) / k;
  *wv = s * 4 * M_PI * (R * cos(kR) - sin(kR) / k) / (k * k);
}

static int grid_init(grid *g, int nx, int ny, int nz, const double *cell, double R, double kT, double mu) {
  memset(g, 0, sizeof *g);
  g->Nx = nx;
  g->Ny = ny;
  g->Nz = nz;
  g->Nkz = nz / 2 + 1;
  g->N = (long long)nx * ny * nz;
  g->Nk = (long long)nx * ny * g->Nkz;
  memcpy(g->cell, cell, sizeof g->cell);
  reciprocal(cell, g->recip);
  g->volume = fabs(cell[0] * (cell[4] * cell[8] - cell[5] * cell[7]) - cell[1] * (cell[3] * cell[8] - cell[5] * cell[6]) +
                   cell[2] * (cell[3] * cell[7] - cell[4] * cell[6]));
  g->dV = g->volume / (double)g->N;
  g->R = R;
  g->kT = kT;
  g->mu = mu;
  g->logNQ = log(nQ(kT));
  if (nz % 2 || !fft_init(&g->px, nx) || !fft_init(&g->py, ny) || !fft_init(&g->pz, nz)) return 0;
  g->w3 = malloc(sizeof(double) * g->Nk);
  g->w2 = malloc(sizeof(double) * g->Nk);
  g->wk = malloc(3 * sizeof(double) * g->Nk);
  for (int f = 0; f < 6; ++f) g->spec[f] = malloc(sizeof(cplx) * g->Nk);
  int maxl = nx > ny ? nx : ny;
  if (nz > maxl) maxl = nz;
  g->line = malloc(sizeof(cplx) * 16 * maxl);
  g->vext = calloc(g->N, sizeof(double));
  const double dr = cbrt(g->dV);
  const double *b = g->recip;
  for (int x = 0; x < nx; ++x)
    for (int y = 0; y < ny; ++y)
      for (int kz = 0; kz < g->Nkz; ++kz) {
        const double mx = wrapi(x, nx), my = wrapi(y, ny), mz = kz;
        const double kx = mx * b[0] + my * b[3] + mz * b[6];
        const double ky = mx * b[1] + my * b[4] + mz * b[7];
        const double kzz = mx * b[2] + my * b[5] + mz * b[8];
        double w3, w2, wv;
        fm_weights(sqrt(kx * kx + ky * ky + kzz * kzz), R, dr, &w3, &w2, &wv);
        const long long i = ((long long)x * ny + y) * g->Nkz + kz;
        g->w3[i] = w3;
        g->w2[i] = w2;
        g->wk[3 * i] = wv * kx;
        g->wk[3 * i + 1] = wv * ky;
        g->wk[3 * i + 2] = wv * kzz;
      }
  return 1;
}

/* r2c of a real field (unnormalised), two z lines per complex FFT */
static void r2c(grid *g, const double *in, cplx *out) {
  const int Nz = g->Nz, Nkz = g->Nkz;
  const long long lines = (long long)g->Nx * g->Ny;
  cplx *z = g->line;
  for (long long l = 0; l < lines; l += 2) {
    const double *a = in + l * Nz;
    const double *bb = (l + 1 < lines) ? in + (l + 1) * Nz : NULL;
    for (int k = 0; k < Nz; ++k) {
      z[k].re = a[k];
      z[k].im = bb ? bb[k] : 0.0;
    }
    fft_run(&g->pz, z, 0);
    for (int k = 0; k < Nkz; ++k) {
      const cplx p = z[k], q = z[(Nz - k) % Nz];
      /* A = (Z_k + conj Z_-k)/2, B = (Z_k - conj Z_-k)/(2i) */
      out[l * Nkz + k].re = 0.5 * (p.re + q.re);
      out[l * Nkz + k].im = 0.5 * (p.im - q.im);
      if (bb) {
        out[(l + 1) * Nkz + k].re = 0.5 * (p.im + q.im);
        out[(l + 1) * Nkz + k].im = -0.5 * (p.re - q.re);
      }
    }
  }
}

/* c2r of a half spectrum (unnormalised), two lines per complex FFT; the
 * imaginary parts of kz = 0 and kz = Nz/2 are dropped */
static void c2r(grid *g, const cplx *in, double *out) {
  const int Nz = g->Nz, Nkz = g->Nkz;
  const long long lines = (long long)g->Nx * g->Ny;
  cplx *z = g->line;
  for (long long l = 0; l < lines; l += 2) {
    const int two = l + 1 < lines;
    for (int k = 0; k < Nz; ++k) {
      cplx A, B = {0, 0};
      if (k < Nkz) {
        A = in[l * Nkz + k];
        if (two) B = in[(l + 1) * Nkz + k];
        if (k == 0 || k == Nz / 2) {
          A.im = 0;
          B.im = 0;
        }
      } else {
        A = in[l * Nkz + (Nz - k)];
        A.im = -A.im;
        if (two) {
          B = in[(l + 1) * Nkz + (Nz - k)];
          B.im = -B.im;
        }
      }
      z[k].re = A.re - B.im;
      z[k].im = A.im + B.re;
    }
    fft_run(&g->pz, z, 1);
    for (int k = 0; k < Nz; ++k) {
      out[l * Nz + k] = z[k].re;
      if (two) out[(l + 1) * Nz + k] = z[k].im;
    }
  }
}

/* FFT along y and along x of a half spectrum, blocks of 8 columns */
static void yx_pass(grid *g, cplx *s, int inverse) {
  const int Nx = g->Nx, Ny = g->Ny, Nkz = g->Nkz;
  enum { B = 8 };
  cplx *t = g->line;
  for (int x = 0; x < Nx; ++x)
    for (int k0 = 0; k0 < Nkz; k0 += B) {
      const int nb = (k0 + B <= Nkz) ? B : Nkz - k0;
      for (int y = 0; y < Ny; ++y)
        for (int c = 0; c < nb; ++c) t[c * Ny + y] = s[((long long)x * Ny + y) * Nkz + k0 + c];
      for (int c = 0; c < nb; ++c) fft_run(&g->py, t + c * Ny, inverse);
      for (int y = 0; y < Ny; ++y)
        for (int c = 0; c < nb; ++c) s[((long long)x * Ny + y) * Nkz + k0 + c] = t[c * Ny + y];
    }
  const long long cols = (long long)Ny * Nkz;
  for (long long c0 = 0; c0 < cols; c0 += B) {
    const int nb = (c0 + B <= cols) ? B : (int)(cols - c0);
    for (int x = 0; x < Nx; ++x)
      for (int c = 0; c < nb; ++c) t[c * Nx + x] = s[(long long)x * cols + c0 + c];
    for (int c = 0; c < nb; ++c) fft_run(&g->px, t + c * Nx, inverse);
    for (int x = 0; x < Nx; ++x)
      for (int c = 0; c < nb; ++c) s[(long long)x * cols + c0 + c] = t[c * Nx + x];
  }
}

static void fwd3(grid *g, const double *in, cplx *out) {
  r2c(g, in, out);
  yx_pass(g, out, 0);
}

static void inv3(grid *g, cplx *in, double *out) {
  yx_pass(g, in, 1);
  c2r(g, in, out);
}

/* ------------------------------------------------------- White-Bear --- */
static void f3_and_derivative(double n, double *f3, double *df3) {
  if (fabs(n) < 0.05) {
    double g = 1.5, gp = 0.0, p = 1.0;
    for (int k = 3; k < 21; ++k) {
      const double c = 2.0 / (k * (k - 1.0) * (k - 2.0));
      gp -= c * (k - 2) * p;
      p *= n;
      g -= c * p;
    }
    const double om = 1.0 - n;
    *f3 = g / (om * om);
    *df3 = gp / (om * om) + 2.0 * g / (om * om * om);
    return;
  }
  const double om = 1.0 - n, L = log1p(-n);
  const double num = n + om * om * L, den = n * n * om * om;
  const double dnum = n - 2.0 * om * L, dden = 2.0 * n * om * (1.0 - 2.0 * n);
  *f3 = num / den;
  *df3 = dnum / den - num * dden / (den * den);
}

static void wb_phi(double n3, double n2, double vx, double vy, double vz, double R, double *p1, double *p2,
                   double *p3, double *d) {
  const double fpR = 4 * M_PI * R, fpR2 = 4 * M_PI * R * R;
  const double omn = 1.0 - n3, L = log1p(-n3), v2 = vx * vx + vy * vy + vz * vz;
  double f3, df3;
  f3_and_derivative(n3, &f3, &df3);
  const double g2 = 1.0 / omn, g2p = 1.0 / (omn * omn);
  const double t3 = n2 * (n2 * n2 - 3.0 * v2);
  *p1 = -n2 * L / fpR2;
  *p2 = (n2 * n2 - v2) * g2 / fpR;
  *p3 = f3 * t3 / (36 * M_PI);
  if (!d) return;
  d[0] = n2 / (fpR2 * omn) + (n2 * n2 - v2) * g2p / fpR + df3 * t3 / (36 * M_PI);
  d[1] = -L / fpR2 + 2.0 * n2 * g2 / fpR + f3 * (3.0 * n2 * n2 - 3.0 * v2) / (36 * M_PI);
  const double cv = -2.0 * g2 / fpR - f3 * 6.0 * n2 / (36 * M_PI);
  d[2] = cv * vx;
  d[3] = cv * vy;
  d[4] = cv * vz;
}

/* Energy (and with grad != NULL the gradient and pgrad = grad/n) of Veff x.
 * Work arrays: nal (5N), n (N). */
typedef struct {
  double *n, *nal, *conv;
} work;

static double evaluate(grid *g, work *w, const double *x, double *grad, double *pgrad, double terms[6]) {
  const double t0 = now();
  const long long N = g->N, Nk = g->Nk;
  const double kT = g->kT;
  double s_ideal = 0, s_ext = 0, s_mun = 0;
  for (long long i = 0; i < N; ++i) {
    const double ve = g->vext[i];
    double nv = 0.0;
    if (!isinf(ve)) {
      nv = exp(-x[i] / kT);
      s_ideal += -nv * (x[i] + kT * g->logNQ + kT);
      if (!isnan(ve)) s_ext += ve * nv;
      s_mun += -g->mu * nv;
    }
    w->n[i] = nv;
  }
  cplx *nk = g->spec[0];
  fwd3(g, w->n, nk);
  const double invN = 1.0 / (double)N;
  for (int f = 0; f < 5; ++f) {
    cplx *o = g->spec[1 + f];
    for (long long i = 0; i < Nk; ++i) {
      const cplx v = nk[i];
      if (f == 0) {
        o[i].re = v.re * g->w3[i] * invN;
        o[i].im = v.im * g->w3[i] * invN;
      } else if (f == 1) {
        o[i].re = v.re * g->w2[i] * invN;
        o[i].im = v.im * g->w2[i] * invN;
      } else {
        /* i wv k_j n~ */
        const double c = g->wk[3 * i + (f - 2)] * invN;
        o[i].re = -v.im * c;
        o[i].im = v.re * c;
      }
    }
    inv3(g, o, w->nal + f * N);
  }
  double s1 = 0, s2 = 0, s3 = 0;
  for (long long i = 0; i < N; ++i) {
    double p1, p2, p3, d[5];
    wb_phi(w->nal[i], w->nal[N + i], w->nal[2 * N + i], w->nal[3 * N + i], w->nal[4 * N + i], g->R, &p1, &p2, &p3,
           grad ? d : NULL);
    s1 += p1;
    s2 += p2;
    s3 += p3;
    if (grad)
      for (int f = 0; f < 5; ++f) w->nal[f * N + i] = d[f];
  }
  const double dV = g->dV;
  terms[0] = kT * s1 * dV;
  terms[1] = kT * s2 * dV;
  terms[2] = kT * s3 * dV;
  terms[3] = s_ideal * dV;
  terms[4] = s_ext * dV;
  terms[5] = s_mun * dV;
  const double E = terms[0] + terms[1] + terms[2] + terms[3] + terms[4] + terms[5];
  if (!grad) {
    g->nevE++;
    g->tE += now() - t0;
    return E;
  }
  /* back-convolution: sum_a w_a(-k) FFT(dPhi/dn_a); vector weights flip sign */
  cplx *acc = g->spec[0];
  for (long long i = 0; i < Nk; ++i) acc[i].re = acc[i].im = 0.0;
  for (int f = 0; f < 5; ++f) {
    cplx *P = g->spec[1 + f];
    fwd3(g, w->nal + f * N, P);
    for (long long i = 0; i < Nk; ++i) {
      const cplx v = P[i];
      if (f == 0) {
        acc[i].re += v.re * g->w3[i] * invN;
        acc[i].im += v.im * g->w3[i] * invN;
      } else if (f == 1) {
        acc[i].re += v.re * g->w2[i] * invN;
        acc[i].im += v.im * g->w2[i] * invN;
      } else {
        const double c = g->wk[3 * i + (f - 2)] * invN;
        acc[i].re -= -v.im * c;
        acc[i].im -= v.re * c;
      }
    }
  }
  inv3(g, acc, w->conv);
  for (long long i = 0; i < N; ++i) {
    const double ve = g->vext[i];
    double gr = 0.0, pg = 0.0;
    if (isfinite(ve)) {
      const double nn = w->n[i];
      const double local = -x[i] - kT * g->logNQ + ve - g->mu;
      const double dfdn = (kT * w->conv[i] + local) * dV;
      gr = -dfdn * nn / kT;
      pg = -dfdn / kT;
    }
    grad[i] = gr;
    pgrad[i] = pg;
  }
  g->nevG++;
  g->tG += now() - t0;
  return E;
}

/* ---------------------------------------------- Minimize (new API) --- */
typedef struct {
  grid *g;
  work *w;
  double *x, *d, *gr, *pg, *og, *xt;
  long long N;
  int iter, maxiter, miniter;
  double precision, relative_precision;
  double step, orig_step, oldgradsqr, deltaE, dEdn, log_avg, error_estimate;
  int have_dir, have_old;
  int E_valid, G_valid;
  double E, terms[6];
  int num_energy, num_grad;
  const char *stop;
} minimize;

static int better(double a, double b) { return a < b || b != b; }

static double m_energy(minimize *m) {
  if (!m->E_valid) {
    m->E = evaluate(m->g, m->w, m->x, NULL, NULL, m->terms);
    m->E_valid = 1;
    m->num_energy++;
  }
  return m->E;
}

/* preconditioned: energy, grad, pgrad from one evaluation */
static void m_pgrad(minimize *m) {
  if (m->G_valid) return;
  m->E = evaluate(m->g, m->w, m->x, m->gr, m->pg, m->terms);
  m->E_valid = 1;
  m->G_valid = 1;
  m->num_energy++;
  m->num_grad++;
}

static void move(minimize *m, double s) {
  for (long long i = 0; i < m->N; ++i) m->x[i] = m->x[i] + s * m->d[i];
}

static void invalidate(minimize *m) { m->E_valid = m->G_valid = 0; }

static double dot(const double *a, const double *b, long long n) {
  double s = 0;
  for (long long i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

/* src/new/Minimize.cpp:9-313 (as oracle/fmt_oracle.py's Minimize) */
static int improve_energy(minimize *m) {
  m->iter++;
  if (m->iter >= m->maxiter) {
    m->stop = "maxiter";
    return 0;
  }
  const double E0 = m_energy(m);
  const double old_deltaE = m->deltaE;
  if (E0 != E0) {
    m->stop = "nan-energy";
    return 0;
  }
  m_pgrad(m);
  invalidate(m);
  const long long N = m->N;
  const double pgg = dot(m->pg, m->gr, N);
  double beta = m->have_old ? (pgg - dot(m->pg, m->og, N)) / m->oldgradsqr : pgg / m->oldgradsqr;
  if (beta < 0 || beta != beta || m->oldgradsqr == 0) beta = 0;
  m->oldgradsqr = pgg;
  for (long long i = 0; i < N; ++i) {
    m->d[i] = m->have_dir ? -m->pg[i] + beta * m->d[i] : -m->pg[i];
    m->og[i] = m->gr[i];
  }
  m->have_dir = m->have_old = 1;
  double gdotd = dot(m->gr, m->d, N);
  if (gdotd > 0) {
    for (long long i = 0; i < N; ++i) m->d[i] = -m->gr[i];
    gdotd = dot(m->gr, m->d, N);
  }
  const double slope = -gdotd;
  if (E0 >= 1.7976931348623157e308 || E0 <= -1.7976931348623157e308) {
    m->stop = "infinite-energy";
    return 0;
  }
  if (slope != slope) {
    m->stop = "nan-slope";
    return 0;
  }
  if (slope == 0) {
    m->stop = "zero-slope";
    return 0;
  }
  if (m->step == m->orig_step && fabs(E0) > 1e-200 && fabs(slope) > 1e-100) m->step = 0.001 * fabs(E0) / slope;
  if (slope * m->step < 0) m->step *= -1;
  double step1 = m->step;
  move(m, step1);
  double Etried = -137, E;
  while (1) {
    step1 *= 0.5;
    move(m, -step1);
    invalidate(m);
    E = m_energy(m);
    if (E == Etried) break;
    Etried = E;
    if (!better(E0, E)) break;
  }
  const double E1 = m_energy(m);
  const double curvature = 2.0 * (E1 - E0 + step1 * slope) / (step1 * step1);
  if (curvature <= 0.0) {
    if (better(E0, E1)) {
      invalidate(m);
      move(m, -step1);
      m->stop = "negative-curvature";
      return 0;
    }
    double Ebest = E0;
    while ((E = m_energy(m)) <= Ebest) {
      Ebest = E;
      move(m, step1);
      invalidate(m);
      step1 *= 2;
    }
    step1 *= 0.5;
    move(m, -step1);
    m->step = step1;
    invalidate(m);
  } else if (E1 == E0) {
    do {
      invalidate(m);
      move(m, -step1);
      step1 *= 2;
      move(m, step1);
      E = m_energy(m);
    } while (E == E0);
    if (E > E0 || E != E) {
      move(m, -step1);
      invalidate(m);
    }
  } else {
    const double step2 = slope / curvature;
    m->step = step2;
    invalidate(m);
    move(m, step2 - step1);
    const double E2 = m_energy(m);
    if (better(E1, E2) && better(E1, E0)) {
      invalidate(m);
      move(m, -(step2 - step1));
      m->step = step1;
    } else if (E2 > E0) {
      invalidate(m);
      move(m, -step2);
      m->stop = "energy-rose";
      return 0;
    }
  }
  const double newE = m_energy(m);
  m->deltaE = newE - E0;
  const double w = 0.01;
  m->dEdn = (1 - w) * m->dEdn + w * m->deltaE;
  m->dEdn = m->deltaE > old_deltaE ? m->deltaE : old_deltaE;
  const double new_log = (m->deltaE != 0 && old_deltaE != 0) ? log(fabs(m->deltaE / old_deltaE)) : 0.0;
  m->log_avg = (1 - w) * m->log_avg + w * new_log;
  const double ratio_avg = exp(m->log_avg);
  double error_guess = fabs((m->dEdn > m->deltaE ? m->dEdn : m->deltaE) / m->log_avg);
  if (ratio_avg >= 1) error_guess = m->precision + fabs(newE);
  m->error_estimate = 2 * error_guess;
  if (m->deltaE == 0 && old_deltaE == 0) {
    m->stop = "no-change";
    return m->iter < m->miniter;
  }
  if (!(m->error_estimate < m->precision) && !(m->error_estimate < m->relative_precision * fabs(newE))) {
    m->stop = "none";
    return 1;
  }
  m->stop = "converged";
  return m->iter < m->miniter;
}

/* --------------------------------------------------------------- main --- */
static int run_eval(const char *inpath, const char *outpath) {
  FILE *in = fopen(inpath, "rb");
  if (!in) return 3;
  char magic[4];
  int32_t n[3];
  double par[12];
  if (fread(magic, 1, 4, in) != 4 || memcmp(magic, "FMTC", 4) || fread(n, 4, 3, in) != 3 ||
      fread(par, 8, 12, in) != 12)
    return 3;
  grid g;
  if (!grid_init(&g, n[0], n[1], n[2], par, par[9], par[10], par[11])) return 4;
  const long long N = g.N;
  double *x = malloc(8 * N), *gr = malloc(8 * N), *pg = malloc(8 * N);
  if (fread(g.vext, 8, N, in) != (size_t)N || fread(x, 8, N, in) != (size_t)N) return 3;
  fclose(in);
  work w = {malloc(8 * N), malloc(5 * 8 * N), malloc(8 * N)};
  double terms[6];
  const double E = evaluate(&g, &w, x, gr, pg, terms);
  FILE *out = fopen(outpath, "wb");
  if (!out) return 3;
  fwrite(&E, 8, 1, out);
  fwrite(terms, 8, 6, out);
  fwrite(gr, 8, N, out);
  fwrite(pg, 8, N, out);
  fclose(out);
  return 0;
}

/* bulk mu of ideal gas + White-Bear (the oracle's chemical_potential) */
static double bulk_mu(double nb, double R, double kT) {
  const double n3 = 4 * M_PI / 3 * R * R * R * nb, n2 = 4 * M_PI * R * R * nb;
  double p1, p2, p3, d[5];
  wb_phi(n3, n2, 0, 0, 0, R, &p1, &p2, &p3, d);
  return kT * log(nb / nQ(kT)) + kT * (d[0] * (4 * M_PI / 3 * R * R * R) + d[1] * (4 * M_PI * R * R));
}

static int run_solute(int argc, char **argv) {
  if (argc < 9) return 2;
  const int n = atoi(argv[2]);
  const double L = atof(argv[3]), eta = atof(argv[4]), radius = atof(argv[5]), precision = atof(argv[6]);
  const int maxiter = atoi(argv[7]);
  const double sample = atof(argv[8]);
  const double R = 1.0, kT = 1.0;
  const double nb = eta / (4 * M_PI / 3 * R * R * R);
  const double cell[9] = {L, 0, 0, 0, L, 0, 0, 0, L};
  grid g;
  if (!grid_init(&g, n, n, n, cell, R, kT, bulk_mu(nb, R, kT))) return 4;
  const long long N = g.N;
  for (int x = 0; x < n; ++x)
    for (int y = 0; y < n; ++y)
      for (int z = 0; z < n; ++z) {
        double r2 = 0;
        const int q[3] = {x, y, z};
        for (int a = 0; a < 3; ++a) {
          double c = q[a] * (L / n);
          c -= L * nearbyint(c / L);
          r2 += c * c;
        }
        if (sqrt(r2) < radius) g.vext[((long long)x * n + y) * n + z] = INFINITY;
      }
  minimize m;
  memset(&m, 0, sizeof m);
  m.g = &g;
  work w = {malloc(8 * N), malloc(5 * 8 * N), malloc(8 * N)};
  m.w = &w;
  m.N = N;
  m.x = malloc(8 * N);
  m.d = calloc(N, 8);
  m.gr = malloc(8 * N);
  m.pg = malloc(8 * N);
  m.og = calloc(N, 8);
  for (long long i = 0; i < N; ++i) m.x[i] = -kT * log(nb);
  m.maxiter = maxiter > 0 ? maxiter : 10000000;
  m.miniter = 6;
  m.precision = precision;
  m.relative_precision = precision > 0 ? 1e-15 : 1e-12;
  m.step = m.orig_step = 0.1;
  m.oldgradsqr = 1.0;
  {
    /* one untimed evaluation first: page in the work arrays and tables */
    double terms[6];
    evaluate(&g, &w, m.x, m.gr, m.pg, terms);
    g.nevE = g.nevG = 0;
    g.tE = g.tG = 0.0;
  }
  const double t0 = now();
  int more = 1;
  while (more) {
    more = improve_energy(&m);
    if (sample > 0 && now() - t0 > sample) break;
  }
  const double dt = now() - t0;
  printf("{\"energy\": %.17g, \"iterations\": %d, \"stop\": \"%s\", \"finished\": %s, \"seconds\": %.6f, "
         "\"energy_evals\": %d, \"grad_evals\": %d, \"seconds_energy\": %.6f, \"seconds_grad\": %.6f}\n",
         m_energy(&m), m.iter, m.stop ? m.stop : "none", more ? "false" : "true", dt, g.nevE, g.nevG, g.tE, g.tG);
  return 0;
}

int main(int argc, char **argv) {
  if (argc >= 4 && !strcmp(argv[1], "eval")) return run_eval(argv[2], argv[3]);
  if (argc >= 2 && !strcmp(argv[1], "solute")) return run_solute(argc, argv);
  fprintf(stderr, "usage: fmt_cpu eval <in> <out> | solute N L eta radius precision maxiter sample_seconds\n");
  return 2;
}
