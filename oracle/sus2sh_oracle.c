/*
 * FP64 CPU oracle for the SUS2-SH potential path.  TEST INFRASTRUCTURE ONLY:
 * tests, the smoke check and the CPU-baseline leg of the benchmark call it as the
 * checker.  It is never part of the engine that is measured or shipped.
 *
 * It restates the reference's "plain" path (all fast-path options off):
 *   .mtp fields                      src/force/sus2_sh.cu:2105-2482 (own minimal parser)
 *   radial tables (LK scaling)       src/force/sus2_sh.cu:2484-2517
 *   Chebyshev-sss radial recurrence  src/force/sus2_sh.cu:3171-3220
 *   real spherical harmonics         src/force/sus2_sh.cu:3032-3169 (here: closed-form solid
 *                                    harmonics expanded into monomials, any l <= 6)
 *   basic moments                    src/force/sus2_sh.cu:3486-3550
 *   flat product graph, energy, reverse  src/force/sus2_sh.cu:4799-4846
 *   edge derivative, forces, virial  src/force/sus2_sh.cu:5996-6035, 6347-6446
 *   multi-image neighbours           src/force/sus2_sh.cu:7657-7767, 9547-9618
 *   parameter gradient dE/dtheta     no reference counterpart (adjoint of the above)
 *   two-layer mu gate                docs/mu-body-order-gate.md:25-63 with the loader fields
 *                                    of src/force/sus2_sh.cu:2153-2266, 2412-2478 and the
 *                                    per-step order of :9821-10256 (gate moments, gate graph,
 *                                    h, 1 + A tanh(a h) per neighbour and channel, gated
 *                                    moments, main graph, forces with the gate adjoint,
 *                                    reverse through the gate graph, gate-layer forces)
 *   ZBL pair term                    src/force/sus2_zbl_common.cuh:171-225 applied per edge as
 *                                    in src/force/sus2_sh.cu:7522-7627; the universal screened
 *                                    Coulomb form of Ziegler, Biersack and Littmark (1985)
 *                                    with the reference's screening constant 2.134563
 * Everything is double precision end to end (the reference keeps float force/virial
 * buffers even in its double mode, src/force/sus2_sh.cuh:200-202).
 *
 * Parity status of each part is recorded in the test-suite: every function here is
 * pinned by closed forms / invariants / finite differences (tests/test_oracle_*.py).
 */
#include <ctype.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_MAX_L 6
#define OR_MAX_COMP ((OR_MAX_L + 1) * (OR_MAX_L + 1))

typedef struct {
  int S, L, K, body, rb, Mr, A, B, Q, T;
  int kind; /* 0 tanh map, 1 rational map */
  double scaling, rc;
  double *shift, *species, *scal, *moment_coeffs, *radial, *type_coeffs;
  int *mapping, *basic_k, *basic_l, *basic_m;
  int *p_left, *p_right, *p_target;
  double *p_coeff;
  /* ZBL: zbl_enabled in the file; per-species Z and per-pair cutoffs set by the caller */
  int zbl, zbl_ready;
  int zbl_z[16];
  double zbl_inner[256], zbl_outer[256];
  /* two-layer mu gate: h_{a,mu} = sum_q weff[mu][q] s_q(a) with weff = the mu-scalar-full
   * weights, or c_{mu,body(q)} w_q for mu-body-linear-combo; G_{ij,mu} = 1 + A tanh(a_tj h) */
  int gate, gate_mode, gate_Qg;
  double gate_A;
  double *gate_radial, *gate_add, *gate_weff; /* [Mr rb], [S], [Mr Qg] */
  int* gate_node;                             /* [Qg] graph node of gate scalar q */
} oracle_model;

/* ------------------------------------------------------------------------ */
/* minimal .mtp reader (same token semantics as sus2_sh.cu:269-477)          */
/* ------------------------------------------------------------------------ */
static char* or_read_file(const char* path)
{
  FILE* f = fopen(path, "rb");
  if (!f) return NULL;
  fseek(f, 0, SEEK_END);
  long n = ftell(f);
  fseek(f, 0, SEEK_SET);
  char* buf = (char*)malloc((size_t)n + 1);
  if (!buf) {
    fclose(f);
    return NULL;
  }
  size_t got = fread(buf, 1, (size_t)n, f);
  buf[got] = 0;
  fclose(f);
  return buf;
}

static const char* or_after(const char* text, const char* token)
{
  const char* p = strstr(text, token);
  return p ? p + strlen(token) : NULL;
}

static int or_int(const char* text, const char* token, int* ok)
{
  const char* p = or_after(text, token);
  if (!p) {
    *ok = 0;
    return 0;
  }
  return (int)strtol(p, NULL, 10);
}

static double or_double(const char* text, const char* token, int* ok)
{
  const char* p = or_after(text, token);
  if (!p) {
    *ok = 0;
    return 0.0;
  }
  return strtod(p, NULL);
}

/* copy of the word after token (trimmed, up to end of line) */
static void or_word(const char* text, const char* token, char* out, int cap, int* ok)
{
  const char* p = or_after(text, token);
  out[0] = 0;
  if (!p) {
    *ok = 0;
    return;
  }
  while (*p == ' ' || *p == '\t') ++p;
  int n = 0;
  while (p[n] && p[n] != '\n' && p[n] != '\r' && n < cap - 1) {
    out[n] = p[n];
    ++n;
  }
  while (n > 0 && isspace((unsigned char)out[n - 1])) --n;
  out[n] = 0;
}

/* [begin, end) of the brace block that starts at the first '{' after token */
static int or_block(const char* text, const char* token, const char** b, const char** e)
{
  const char* p = strstr(text, token);
  if (!p) return 0;
  p = strchr(p, '{');
  if (!p) return 0;
  int depth = 0;
  for (const char* q = p; *q; ++q) {
    if (*q == '{') ++depth;
    if (*q == '}' && --depth == 0) {
      *b = p;
      *e = q + 1;
      return 1;
    }
  }
  return 0;
}

/* parse all numbers in [b, e) treating braces and commas as blanks */
static int or_numbers(const char* b, const char* e, double* out, int cap)
{
  int n = 0;
  const char* p = b;
  while (p < e) {
    if (*p == '{' || *p == '}' || *p == ',' || isspace((unsigned char)*p)) {
      ++p;
      continue;
    }
    char* end = NULL;
    double v = strtod(p, &end);
    if (end == p) break;
    if (n < cap) out[n] = v;
    ++n;
    p = end;
  }
  return n;
}

static double* or_array(const char* text, const char* token, int expect, int* ok)
{
  const char *b, *e;
  if (!or_block(text, token, &b, &e)) {
    *ok = 0;
    return NULL;
  }
  double* v = (double*)calloc((size_t)(expect > 0 ? expect : 1), sizeof(double));
  int n = or_numbers(b, e, v, expect);
  if (n != expect) *ok = 0;
  return v;
}

void oracle_free(oracle_model* m)
{
  if (!m) return;
  free(m->shift);
  free(m->species);
  free(m->scal);
  free(m->moment_coeffs);
  free(m->radial);
  free(m->type_coeffs);
  free(m->mapping);
  free(m->basic_k);
  free(m->basic_l);
  free(m->basic_m);
  free(m->p_left);
  free(m->p_right);
  free(m->p_target);
  free(m->p_coeff);
  free(m->gate_radial);
  free(m->gate_add);
  free(m->gate_weff);
  free(m->gate_node);
  free(m);
}

/* Gate fields of a two_layer_gate_enabled model (only the accepted form: mu-body-linear-combo
 * or mu-scalar-full, neighbour site, shared radial scaling; the engine's loader enforces the
 * reference's rejections).  Returns 0 on a parse problem. */
static int or_load_gate(const char* text, oracle_model* m)
{
  int ok = 1;
  char word[128];
  or_word(text, "two_layer_gate_mode =", word, sizeof word, &ok);
  if (!ok) return 0;
  if (strcmp(word, "mu-body-linear-combo") == 0)
    m->gate_mode = 1;
  else if (strcmp(word, "mu-scalar-full") == 0 || strcmp(word, "mu-body-order") == 0)
    m->gate_mode = 2;
  else
    return 0;
  m->gate_A = strstr(text, "two_layer_gate_tanh_amplitude =")
                ? or_double(text, "two_layer_gate_tanh_amplitude =", &ok)
                : 0.8;
  const int Mr = m->Mr, S = m->S, K = m->K;
  m->gate_radial = or_array(text, "two_layer_gate_radial_coeffs =", Mr * m->rb, &ok);
  if (strstr(text, "two_layer_gate_additive_coeffs ="))
    m->gate_add = or_array(text, "two_layer_gate_additive_coeffs =", S, &ok);
  else {
    m->gate_add = (double*)calloc((size_t)S, sizeof(double));
    for (int t = 0; t < S; ++t) m->gate_add[t] = 1.0;
  }
  const int Qg = or_int(text, "two_layer_gate_weight_count =", &ok) / (m->gate_mode == 2 ? Mr : 1);
  if (!ok || Qg <= 0) return 0;
  m->gate_Qg = Qg;
  double* idx = or_array(text, "two_layer_gate_scalar_indices =", Qg, &ok);
  double* w = or_array(text, "two_layer_gate_weights =", m->gate_mode == 2 ? Mr * Qg : Qg, &ok);
  double* info = or_array(text, "sh_scalar_info =", m->Q, &ok);
  double* mix = m->gate_mode == 1 ? or_array(text, "two_layer_gate_body_mix_weights =", Mr * K, &ok)
                                  : NULL;
  m->gate_node = (int*)calloc((size_t)Qg, sizeof(int));
  m->gate_weff = (double*)calloc((size_t)(Mr * Qg), sizeof(double));
  for (int q = 0; ok && q < Qg; ++q) {
    const int s = (int)idx[q];
    if (s < 0 || s >= m->Q) {
      ok = 0;
      break;
    }
    m->gate_node[q] = m->mapping[s];
    const int body = (int)info[s] - 2;
    if (body < 0 || body >= K) {
      ok = 0;
      break;
    }
    for (int mu = 0; mu < Mr; ++mu)
      m->gate_weff[mu * Qg + q] = m->gate_mode == 2 ? w[mu * Qg + q] : mix[mu * K + body] * w[q];
  }
  free(idx);
  free(w);
  free(info);
  free(mix);
  return ok;
}

/* Returns NULL on any parse problem (the oracle only reads files it is given by tests). */
oracle_model* oracle_load(const char* path)
{
  char* text = or_read_file(path);
  if (!text) return NULL;
  int ok = 1;
  oracle_model* m = (oracle_model*)calloc(1, sizeof(oracle_model));
  char word[128];
  or_word(text, "potential_tag =", word, sizeof word, &ok);
  if (strcmp(word, "SUS2-SH") != 0) ok = 0;
  or_word(text, "radial_basis_type =", word, sizeof word, &ok);
  if (strncmp(word, "RBLaguerre_log1p", 16) == 0)
    m->kind = 2;
  else if (strncmp(word, "RBChebyshev_sss_rational", 24) == 0)
    m->kind = 1;
  else if (strncmp(word, "RBChebyshev_sss", 15) == 0)
    m->kind = 0;
  else
    ok = 0;
  m->scaling = or_double(text, "scaling =", &ok);
  m->S = or_int(text, "species_count =", &ok);
  m->L = or_int(text, "sh_l_max =", &ok);
  m->K = or_int(text, "sh_k_max =", &ok);
  m->body = or_int(text, "sh_body_order =", &ok);
  m->rc = or_double(text, "max_dist =", &ok);
  m->rb = or_int(text, "radial_basis_size =", &ok);
  m->Mr = or_int(text, "radial_funcs_count =", &ok);
  m->A = or_int(text, "alpha_moments_count =", &ok);
  m->B = or_int(text, "alpha_index_basic_count =", &ok);
  m->Q = or_int(text, "alpha_scalar_moments =", &ok);
  m->T = or_int(text, "sh_product_count =", &ok);
  {
    char zw[32];
    int zok = 1;
    zw[0] = 0;
    if (strstr(text, "zbl_enabled =")) or_word(text, "zbl_enabled =", zw, sizeof zw, &zok);
    m->zbl = zok && (strcmp(zw, "1") == 0 || strcmp(zw, "true") == 0 || strcmp(zw, "TRUE") == 0 ||
                     strcmp(zw, "yes") == 0 || strcmp(zw, "on") == 0);
  }
  if (!ok || m->S <= 0 || m->L < 0 || m->L > OR_MAX_L || m->K <= 0 || m->Mr != m->K * (m->L + 1) ||
      m->B <= 0 || m->A < m->B || m->T < 0 || m->Q <= 0 || m->rb <= 0) {
    free(text);
    oracle_free(m);
    return NULL;
  }
  const int S = m->S;
  m->shift = or_array(text, "shift_coeffs =", S, &ok);
  m->species = or_array(text, "species_coeffs =", S, &ok);
  m->scal = or_array(text, "scal_coeffs =", 2 * S * S * m->Mr, &ok);
  m->moment_coeffs = or_array(text, "moment_coeffs =", m->Q, &ok);
  double* map = or_array(text, "alpha_moment_mapping =", m->Q, &ok);
  double* bas = or_array(text, "alpha_index_basic =", 3 * m->B, &ok);
  double* prod = or_array(text, "sh_products =", 4 * m->T, &ok);
  m->mapping = (int*)calloc((size_t)m->Q, sizeof(int));
  m->basic_k = (int*)calloc((size_t)m->B, sizeof(int));
  m->basic_l = (int*)calloc((size_t)m->B, sizeof(int));
  m->basic_m = (int*)calloc((size_t)m->B, sizeof(int));
  m->p_left = (int*)calloc((size_t)(m->T + 1), sizeof(int));
  m->p_right = (int*)calloc((size_t)(m->T + 1), sizeof(int));
  m->p_target = (int*)calloc((size_t)(m->T + 1), sizeof(int));
  m->p_coeff = (double*)calloc((size_t)(m->T + 1), sizeof(double));
  if (ok) {
    for (int s = 0; s < m->Q; ++s) m->mapping[s] = (int)map[s];
    for (int b = 0; b < m->B; ++b) {
      m->basic_k[b] = (int)bas[3 * b];
      m->basic_l[b] = (int)bas[3 * b + 1];
      m->basic_m[b] = (int)bas[3 * b + 2];
    }
    for (int p = 0; p < m->T; ++p) {
      m->p_left[p] = (int)prod[4 * p];
      m->p_right[p] = (int)prod[4 * p + 1];
      m->p_target[p] = (int)prod[4 * p + 2];
      m->p_coeff[p] = prod[4 * p + 3];
    }
  }
  free(map);
  free(bas);
  free(prod);
  /* radial rows: first Mr brace groups after "radial_coeffs" (sus2_sh.cu:2380-2402) */
  m->radial = (double*)calloc((size_t)(m->Mr * m->rb), sizeof(double));
  m->type_coeffs = (double*)calloc((size_t)S, sizeof(double));
  const char* p = strstr(text, "radial_coeffs");
  const char* stop = p ? strstr(p, "alpha_moments_count") : NULL;
  if (!p || !stop) ok = 0;
  int row = 0;
  double vals[64];
  while (ok && row < m->Mr) {
    const char* b = strchr(p, '{');
    if (!b || b >= stop) {
      ok = 0;
      break;
    }
    const char* e = strchr(b, '}');
    if (!e) {
      ok = 0;
      break;
    }
    int n = or_numbers(b, e + 1, vals, 64);
    if (n != m->rb + S) {
      ok = 0;
      break;
    }
    for (int k = 0; k < m->rb; ++k) m->radial[row * m->rb + k] = vals[k];
    if (row == 0)
      for (int t = 0; t < S; ++t) m->type_coeffs[t] = vals[m->rb + t];
    ++row;
    p = e + 1;
  }
  {
    char gw[32];
    int gok = 1;
    gw[0] = 0;
    if (strstr(text, "two_layer_gate_enabled ="))
      or_word(text, "two_layer_gate_enabled =", gw, sizeof gw, &gok);
    m->gate = gok && (strcmp(gw, "1") == 0 || strcmp(gw, "true") == 0 || strcmp(gw, "TRUE") == 0 ||
                      strcmp(gw, "yes") == 0 || strcmp(gw, "on") == 0);
    if (ok && m->gate && !or_load_gate(text, m)) ok = 0;
  }
  free(text);
  if (!ok) {
    oracle_free(m);
    return NULL;
  }
  return m;
}

void oracle_dims(const oracle_model* m, int* out /* 12 */)
{
  out[0] = m->S;
  out[1] = m->L;
  out[2] = m->K;
  out[3] = m->body;
  out[4] = m->rb;
  out[5] = m->Mr;
  out[6] = m->A;
  out[7] = m->B;
  out[8] = m->Q;
  out[9] = m->T;
  out[10] = m->kind;
  out[11] = m->gate ? m->gate_Qg : 0;
}

double oracle_cutoff(const oracle_model* m) { return m->rc; }

/* ------------------------------------------------------------------------ */
/* real solid harmonics as monomial tables                                   */
/* Y_lm(r) = sqrt((2l+1)/4pi) * f_m * Pi_l^|m|(z,r) * {A_m | B_|m|}(x,y) / r^l */
/* (standard closed form; sign convention of sus2_sh.cu:3058-3168)           */
/* ------------------------------------------------------------------------ */
typedef struct {
  int n;
  int a[64], b[64], c[64];
  double coef[64];
} or_poly;

static or_poly or_ylm[OR_MAX_COMP];
static int or_ylm_ready = 0;

static double or_fact(int n)
{
  double v = 1.0;
  for (int i = 2; i <= n; ++i) v *= i;
  return v;
}

static double or_binom(int n, int k)
{
  if (k < 0 || k > n) return 0.0;
  return or_fact(n) / (or_fact(k) * or_fact(n - k));
}

static void or_poly_add(double grid[OR_MAX_L + 1][OR_MAX_L + 1][OR_MAX_L + 1], int a, int b, int c,
                        double v)
{
  grid[a][b][c] += v;
}

static void or_init_ylm(void)
{
  if (or_ylm_ready) return;
  const double pi = 3.14159265358979323846264338327950288;
  for (int l = 0; l <= OR_MAX_L; ++l) {
    for (int m = -l; m <= l; ++m) {
      const int am = m < 0 ? -m : m;
      double grid[OR_MAX_L + 1][OR_MAX_L + 1][OR_MAX_L + 1];
      memset(grid, 0, sizeof grid);
      const double norm = sqrt((2.0 * l + 1.0) / (4.0 * pi)) * (m == 0 ? 1.0 : sqrt(2.0)) *
                          sqrt(or_fact(l - am) / or_fact(l + am));
      /* azimuthal factor: A_m (m>0, cos) or B_|m| (m<0, sin) */
      for (int p = 0; p <= am; ++p) {
        double trig;
        const int q = (am - p) & 3; /* angle (am-p) pi/2 */
        if (m >= 0)
          trig = (q == 0) ? 1.0 : (q == 2 ? -1.0 : 0.0);
        else
          trig = (q == 1) ? 1.0 : (q == 3 ? -1.0 : 0.0);
        if (trig == 0.0) continue;
        const double az = or_binom(am, p) * trig; /* x^p y^(am-p) */
        for (int k = 0; 2 * k <= l - am; ++k) {
          const double pk = ((k & 1) ? -1.0 : 1.0) * pow(2.0, -l) * or_binom(l, k) *
                            or_binom(2 * l - 2 * k, l) * or_fact(l - 2 * k) /
                            or_fact(l - 2 * k - am);
          /* r^(2k) = sum_{i+j+h=k} k!/(i!j!h!) x^2i y^2j z^2h, times z^(l-2k-am) */
          for (int i = 0; i <= k; ++i)
            for (int j = 0; i + j <= k; ++j) {
              const int h = k - i - j;
              const double mult = or_fact(k) / (or_fact(i) * or_fact(j) * or_fact(h));
              or_poly_add(grid, p + 2 * i, am - p + 2 * j, l - 2 * k - am + 2 * h,
                          norm * az * pk * mult);
            }
        }
      }
      or_poly* P = &or_ylm[l * l + m + l];
      P->n = 0;
      for (int a = 0; a <= l; ++a)
        for (int b = 0; a + b <= l; ++b) {
          const int c = l - a - b;
          if (grid[a][b][c] != 0.0 && fabs(grid[a][b][c]) > 1e-15) {
            P->a[P->n] = a;
            P->b[P->n] = b;
            P->c[P->n] = c;
            P->coef[P->n] = grid[a][b][c];
            P->n++;
          }
        }
    }
  }
  or_ylm_ready = 1;
}

static double or_ipow(double x, int n)
{
  double v = 1.0;
  for (int i = 0; i < n; ++i) v *= x;
  return v;
}

/* values Y[y] and gradients dY[3y+c] with respect to the displacement vector */
void oracle_ylm(int lmax, double x, double y, double z, double* Y, double* dY)
{
  or_init_ylm();
  const double r2 = x * x + y * y + z * z;
  const double r = sqrt(r2);
  for (int l = 0; l <= lmax; ++l) {
    const double inv_rl = 1.0 / or_ipow(r, l);
    for (int m = -l; m <= l; ++m) {
      const int yi = l * l + m + l;
      const or_poly* P = &or_ylm[yi];
      double v = 0.0, gx = 0.0, gy = 0.0, gz = 0.0;
      for (int t = 0; t < P->n; ++t) {
        const int a = P->a[t], b = P->b[t], c = P->c[t];
        const double cf = P->coef[t];
        const double xa = or_ipow(x, a), yb = or_ipow(y, b), zc = or_ipow(z, c);
        v += cf * xa * yb * zc;
        if (a > 0) gx += cf * a * or_ipow(x, a - 1) * yb * zc;
        if (b > 0) gy += cf * b * xa * or_ipow(y, b - 1) * zc;
        if (c > 0) gz += cf * c * xa * yb * or_ipow(z, c - 1);
      }
      Y[yi] = v * inv_rl;
      if (dY) {
        /* d(P/r^l) = dP/r^l - l P r_vec / r^(l+2) */
        const double f = (double)l * v * inv_rl / r2;
        dY[3 * yi + 0] = gx * inv_rl - f * x;
        dY[3 * yi + 1] = gy * inv_rl - f * y;
        dY[3 * yi + 2] = gz * inv_rl - f * z;
      }
    }
  }
}

/* ------------------------------------------------------------------------ */
/* radial functions (sus2_sh.cu:2484-2517, 3171-3220)                        */
/* ------------------------------------------------------------------------ */
/* R[mu], dR[mu] for pair (ti, tj) at distance r.  Optionally also the pieces needed
   for the parameter adjoint: T[mu*rb+n] = (r-rc)^2 T_n(xi), and dxi/da, dxi/ds.   */
static void or_radial_full(const oracle_model* m, int ti, int tj, double r, double* R, double* dR,
                           double* Tn, double* dTdxi, double* xi_out, double* dxi_da,
                           double* dxi_ds)
{
  const int S = m->S;
  const double dr = r - m->rc;
  const double env = dr * dr, denv = 2.0 * dr;
  const double tscale = m->scaling * m->type_coeffs[ti] * m->type_coeffs[tj];
  double T[64], dT[64];
  if (m->kind == 2) {
    /* RBLaguerre_log1p (sus2_sh.cu:3299-3354): phi_n = env e^(-u/2) L_n(u),
     * u = a log(1 + r / rho), rho = max(shift, 1e-8); L_n by the three-term recurrence and
     * L_n' from L_{n+1}' = L_n' - L_n (dE/dtheta of (a, rho) is not provided) */
    for (int mu = 0; mu < m->Mr; ++mu) {
      const int so = 2 * mu * S * S + S * ti + tj;
      const double a = m->scal[so];
      const double rho = m->scal[so + S * S] > 1e-8 ? m->scal[so + S * S] : 1e-8;
      const double u = a * log1p(r / rho), du = a / (rho + r), ex = exp(-0.5 * u);
      double Lg[64], dL[64];
      Lg[0] = 1.0;
      dL[0] = 0.0;
      if (m->rb > 1) {
        Lg[1] = 1.0 - u;
        dL[1] = -1.0;
      }
      for (int n = 1; n + 1 < m->rb; ++n) {
        Lg[n + 1] = ((2.0 * n + 1.0 - u) * Lg[n] - n * Lg[n - 1]) / (n + 1.0);
        dL[n + 1] = dL[n] - Lg[n];
      }
      double acc = 0.0, accd = 0.0;
      for (int n = 0; n < m->rb; ++n) {
        const double c = m->radial[mu * m->rb + n] * tscale;
        const double phi = env * ex * Lg[n];
        const double dphi = ex * (denv * Lg[n] + env * du * (dL[n] - 0.5 * Lg[n]));
        acc += c * phi;
        accd += c * dphi;
        if (Tn) Tn[mu * m->rb + n] = phi;
        if (dTdxi) dTdxi[mu * m->rb + n] = 0.0;
      }
      R[mu] = acc;
      if (dR) dR[mu] = accd;
      if (xi_out) xi_out[mu] = 0.0;
      if (dxi_da) dxi_da[mu] = 0.0;
      if (dxi_ds) dxi_ds[mu] = 0.0;
    }
    return;
  }
  for (int mu = 0; mu < m->Mr; ++mu) {
    const int so = 2 * mu * S * S + S * ti + tj;
    const double a = m->scal[so];
    const double s = m->scal[so + S * S];
    const double zz = 0.5 * a * (r - s);
    double xi, dxi_dz;
    if (m->kind == 1) {
      const double q = 1.0 / sqrt(1.0 + zz * zz);
      xi = zz * q;
      dxi_dz = q * q * q;
    } else {
      xi = tanh(zz);
      dxi_dz = 1.0 - xi * xi;
    }
    const double dxi_dr = 0.5 * a * dxi_dz;
    /* Chebyshev recurrence T_n and dT_n/dxi */
    T[0] = 1.0;
    dT[0] = 0.0;
    if (m->rb > 1) {
      T[1] = xi;
      dT[1] = 1.0;
    }
    for (int n = 2; n < m->rb; ++n) {
      T[n] = 2.0 * xi * T[n - 1] - T[n - 2];
      dT[n] = 2.0 * (T[n - 1] + xi * dT[n - 1]) - dT[n - 2];
    }
    double acc = 0.0, accd = 0.0;
    for (int n = 0; n < m->rb; ++n) {
      const double c = m->radial[mu * m->rb + n] * tscale;
      acc += c * T[n];
      accd += c * dT[n];
      if (Tn) Tn[mu * m->rb + n] = env * T[n];
      if (dTdxi) dTdxi[mu * m->rb + n] = env * dT[n];
    }
    R[mu] = env * acc;
    if (dR) dR[mu] = denv * acc + env * dxi_dr * accd;
    if (xi_out) xi_out[mu] = xi;
    if (dxi_da) dxi_da[mu] = 0.5 * (r - s) * dxi_dz;
    if (dxi_ds) dxi_ds[mu] = -0.5 * a * dxi_dz;
  }
}

void oracle_radial(const oracle_model* m, int ti, int tj, double r, double* R, double* dR)
{
  or_radial_full(m, ti, tj, r, R, dR, NULL, NULL, NULL, NULL, NULL);
}

/* ------------------------------------------------------------------------ */
/* neighbours: every periodic image within rc (sus2_sh.cu:7668-7767)         */
/* ------------------------------------------------------------------------ */
typedef struct {
  int* start; /* N+1 */
  int* j;
  double* d; /* 3 per edge: r_j(image) - r_i */
} or_nlist;

static void or_free_nlist(or_nlist* nl)
{
  free(nl->start);
  free(nl->j);
  free(nl->d);
}

static void or_box_thickness(const double* h, double* t)
{
  /* columns a=(h0,h3,h6), b=(h1,h4,h7), c=(h2,h5,h8) */
  const double a[3] = {h[0], h[3], h[6]}, b[3] = {h[1], h[4], h[7]}, c[3] = {h[2], h[5], h[8]};
  double bc[3] = {b[1] * c[2] - b[2] * c[1], b[2] * c[0] - b[0] * c[2], b[0] * c[1] - b[1] * c[0]};
  double ca[3] = {c[1] * a[2] - c[2] * a[1], c[2] * a[0] - c[0] * a[2], c[0] * a[1] - c[1] * a[0]};
  double ab[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
  const double vol = fabs(a[0] * bc[0] + a[1] * bc[1] + a[2] * bc[2]);
  t[0] = vol / sqrt(bc[0] * bc[0] + bc[1] * bc[1] + bc[2] * bc[2]);
  t[1] = vol / sqrt(ca[0] * ca[0] + ca[1] * ca[1] + ca[2] * ca[2]);
  t[2] = vol / sqrt(ab[0] * ab[0] + ab[1] * ab[1] + ab[2] * ab[2]);
}

/* Brute force over all j and all images within range (reference O(N^2) path). */
static int or_neighbors_brute(int N, const double* pos, const double* h, const int* pbc, double rc,
                              or_nlist* nl)
{
  double t[3];
  or_box_thickness(h, t);
  int R[3];
  for (int d = 0; d < 3; ++d) R[d] = pbc[d] ? (int)ceil(rc / t[d]) : 0;
  const double rc2 = rc * rc;
  int* cnt = (int*)calloc((size_t)N + 1, sizeof(int));
  for (int pass = 0; pass < 2; ++pass) {
    if (pass == 1) {
      nl->start = (int*)malloc(((size_t)N + 1) * sizeof(int));
      nl->start[0] = 0;
      for (int i = 0; i < N; ++i) nl->start[i + 1] = nl->start[i] + cnt[i];
      nl->j = (int*)malloc(((size_t)nl->start[N] + 1) * sizeof(int));
      nl->d = (double*)malloc(((size_t)nl->start[N] + 1) * 3 * sizeof(double));
    }
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < N; ++i) {
      int c = 0;
      for (int j = 0; j < N; ++j) {
        const double d0[3] = {pos[j] - pos[i], pos[N + j] - pos[N + i],
                              pos[2 * N + j] - pos[2 * N + i]};
        for (int sx = -R[0]; sx <= R[0]; ++sx)
          for (int sy = -R[1]; sy <= R[1]; ++sy)
            for (int sz = -R[2]; sz <= R[2]; ++sz) {
              if (j == i && sx == 0 && sy == 0 && sz == 0) continue;
              const double dx = d0[0] + sx * h[0] + sy * h[1] + sz * h[2];
              const double dy = d0[1] + sx * h[3] + sy * h[4] + sz * h[5];
              const double dz = d0[2] + sx * h[6] + sy * h[7] + sz * h[8];
              if (dx * dx + dy * dy + dz * dz < rc2) {
                if (pass == 1) {
                  const int e = nl->start[i] + c;
                  nl->j[e] = j;
                  nl->d[3 * e] = dx;
                  nl->d[3 * e + 1] = dy;
                  nl->d[3 * e + 2] = dz;
                }
                ++c;
              }
            }
      }
      if (pass == 0) cnt[i] = c;
    }
  }
  free(cnt);
  return 0;
}

/* Cell-list search over all images within rc (for large systems). */
static int or_neighbors_cells(int N, const double* pos, const double* h, const int* pbc, double rc,
                              or_nlist* nl)
{
  double t[3];
  or_box_thickness(h, t);
  const double* hi = h + 9; /* inverse, row-major */
  int n[3], R[3];
  for (int d = 0; d < 3; ++d) {
    n[d] = (int)floor(t[d] / rc);
    if (n[d] < 1) n[d] = 1;
    if (n[d] > 512) n[d] = 512;
    R[d] = (int)ceil(rc * n[d] / t[d]);
    if (R[d] < 1) R[d] = 1;
  }
  const long ncell = (long)n[0] * n[1] * n[2];
  double* w = (double*)malloc((size_t)N * 3 * sizeof(double)); /* wrapped positions */
  int* cell = (int*)malloc((size_t)N * sizeof(int));
  int* head = (int*)malloc((size_t)(ncell + 1) * sizeof(int));
  int* order = (int*)malloc((size_t)N * sizeof(int));
  for (long c = 0; c <= ncell; ++c) head[c] = 0;
  for (int i = 0; i < N; ++i) {
    const double x = pos[i], y = pos[N + i], z = pos[2 * N + i];
    double s[3] = {hi[0] * x + hi[1] * y + hi[2] * z, hi[3] * x + hi[4] * y + hi[5] * z,
                   hi[6] * x + hi[7] * y + hi[8] * z};
    int ci[3];
    for (int d = 0; d < 3; ++d) {
      if (pbc[d]) {
        s[d] -= floor(s[d]);
        if (s[d] >= 1.0) s[d] = 0.0;
      }
      int c = (int)floor(s[d] * n[d]);
      if (c < 0) c = 0;
      if (c >= n[d]) c = n[d] - 1;
      ci[d] = c;
    }
    w[3 * i] = h[0] * s[0] + h[1] * s[1] + h[2] * s[2];
    w[3 * i + 1] = h[3] * s[0] + h[4] * s[1] + h[5] * s[2];
    w[3 * i + 2] = h[6] * s[0] + h[7] * s[1] + h[8] * s[2];
    cell[i] = (ci[2] * n[1] + ci[1]) * n[0] + ci[0];
    head[cell[i] + 1]++;
  }
  for (long c = 0; c < ncell; ++c) head[c + 1] += head[c];
  {
    int* fill = (int*)malloc((size_t)ncell * sizeof(int));
    for (long c = 0; c < ncell; ++c) fill[c] = head[c];
    for (int i = 0; i < N; ++i) order[fill[cell[i]]++] = i;
    free(fill);
  }
  const double rc2 = rc * rc;
  int* cnt = (int*)calloc((size_t)N + 1, sizeof(int));
  for (int pass = 0; pass < 2; ++pass) {
    if (pass == 1) {
      nl->start = (int*)malloc(((size_t)N + 1) * sizeof(int));
      nl->start[0] = 0;
      for (int i = 0; i < N; ++i) nl->start[i + 1] = nl->start[i] + cnt[i];
      nl->j = (int*)malloc(((size_t)nl->start[N] + 1) * sizeof(int));
      nl->d = (double*)malloc(((size_t)nl->start[N] + 1) * 3 * sizeof(double));
    }
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < N; ++i) {
      const int c0 = cell[i];
      const int cx = c0 % n[0], cy = (c0 / n[0]) % n[1], cz = c0 / (n[0] * n[1]);
      int c = 0;
      for (int oz = -R[2]; oz <= R[2]; ++oz) {
        int z = cz + oz, sz = 0;
        if (pbc[2]) {
          sz = (int)floor((double)z / n[2]);
          z -= sz * n[2];
        } else if (z < 0 || z >= n[2]) {
          continue;
        }
        for (int oy = -R[1]; oy <= R[1]; ++oy) {
          int y = cy + oy, sy = 0;
          if (pbc[1]) {
            sy = (int)floor((double)y / n[1]);
            y -= sy * n[1];
          } else if (y < 0 || y >= n[1]) {
            continue;
          }
          for (int ox = -R[0]; ox <= R[0]; ++ox) {
            int x = cx + ox, sx = 0;
            if (pbc[0]) {
              sx = (int)floor((double)x / n[0]);
              x -= sx * n[0];
            } else if (x < 0 || x >= n[0]) {
              continue;
            }
            const int cc = (z * n[1] + y) * n[0] + x;
            const double shx = sx * h[0] + sy * h[1] + sz * h[2];
            const double shy = sx * h[3] + sy * h[4] + sz * h[5];
            const double shz = sx * h[6] + sy * h[7] + sz * h[8];
            for (int k = head[cc]; k < head[cc + 1]; ++k) {
              const int j = order[k];
              if (j == i && sx == 0 && sy == 0 && sz == 0) continue;
              const double dx = w[3 * j] - w[3 * i] + shx;
              const double dy = w[3 * j + 1] - w[3 * i + 1] + shy;
              const double dz = w[3 * j + 2] - w[3 * i + 2] + shz;
              if (dx * dx + dy * dy + dz * dz < rc2) {
                if (pass == 1) {
                  const int e = nl->start[i] + c;
                  nl->j[e] = j;
                  nl->d[3 * e] = dx;
                  nl->d[3 * e + 1] = dy;
                  nl->d[3 * e + 2] = dz;
                }
                ++c;
              }
            }
          }
        }
      }
      if (pass == 0) cnt[i] = c;
    }
  }
  free(cnt);
  free(w);
  free(cell);
  free(head);
  free(order);
  return 0;
}

/* neighbour counts for tests: mode 0 brute force, 1 cell list */
int oracle_neighbor_counts(int N, const double* pos, const double* h, const int* pbc, double rc,
                           int mode, int* counts)
{
  or_nlist nl;
  if (mode == 0)
    or_neighbors_brute(N, pos, h, pbc, rc, &nl);
  else
    or_neighbors_cells(N, pos, h, pbc, rc, &nl);
  for (int i = 0; i < N; ++i) counts[i] = nl.start[i + 1] - nl.start[i];
  or_free_nlist(&nl);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* energy / forces / virial / parameter gradient                             */
/* ------------------------------------------------------------------------ */
/* Parameter vector layout (also used by the engine's dE/dtheta):
 *   [shift S][species S][moment_coeffs Q][radial Mr*rb][type_coeffs S][scal 2 S^2 Mr]  */
int oracle_param_count(const oracle_model* m)
{
  return 2 * m->S + m->Q + m->Mr * m->rb + m->S + 2 * m->S * m->S * m->Mr;
}

/* mode bit 0: use brute-force neighbours (else cell list).
 * dparam (optional, length oracle_param_count): accumulates dE_total/dtheta.        */
/* Flat product graph on the moments M[0..A) whose basics M[0..B) are set
 * (src/force/sus2_sh.cu:4799-4846): M[t] += c M[l] M[r] in file order, returns
 * E' = sum_s w_s M[map[s]]. */
static double or_graph_forward(const oracle_model* m, double* M)
{
  for (int p = 0; p < m->T; ++p)
    M[m->p_target[p]] += m->p_coeff[p] * M[m->p_left[p]] * M[m->p_right[p]];
  double sum = 0.0;
  for (int s = 0; s < m->Q; ++s) sum += m->moment_coeffs[s] * M[m->mapping[s]];
  return sum;
}

/* Reverse of or_graph_forward scaled by `seed`: G = seed * dE'/dM for every node. */
static void or_graph_reverse(const oracle_model* m, double seed, const double* M, double* G)
{
  for (int a = 0; a < m->A; ++a) G[a] = 0.0;
  for (int s = 0; s < m->Q; ++s) G[m->mapping[s]] += seed * m->moment_coeffs[s];
  for (int p = m->T - 1; p >= 0; --p) {
    const double g = G[m->p_target[p]] * m->p_coeff[p];
    G[m->p_left[p]] += g * M[m->p_right[p]];
    G[m->p_right[p]] += g * M[m->p_left[p]];
  }
}

/* Graph part alone for `n` atoms given their basic moments basics[n][B]:
 * e_prime[n] = E', grad[n][B] = dE'/d(basic). */
int oracle_graph(const oracle_model* m, int n, const double* basics, double* e_prime, double* grad)
{
  double* M = (double*)malloc((size_t)m->A * sizeof(double));
  double* G = (double*)malloc((size_t)m->A * sizeof(double));
  if (!M || !G) {
    free(M);
    free(G);
    return -1;
  }
  for (int i = 0; i < n; ++i) {
    for (int a = 0; a < m->A; ++a) M[a] = a < m->B ? basics[(size_t)i * m->B + a] : 0.0;
    e_prime[i] = or_graph_forward(m, M);
    or_graph_reverse(m, 1.0, M, G);
    for (int b = 0; b < m->B; ++b) grad[(size_t)i * m->B + b] = G[b];
  }
  free(M);
  free(G);
  return 0;
}

/* ZBL pair energy phi(r) and d phi / dr: screened Coulomb 14.3996454784255 Zi Zj / r times
 * sum_k a_k exp(-b_k r / a), 1 / a = 2.134563 (Zi^0.23 + Zj^0.23), times the cosine switch
 * (1 + cos(pi (r - inner) / (outer - inner))) / 2 above inner; zero from outer on. */
void oracle_zbl_pair(int zi, int zj, double r, double inner, double outer, double* e, double* de)
{
  static const double a[4] = {0.18175, 0.50986, 0.28022, 0.02817};
  static const double b[4] = {3.1998, 0.94229, 0.4029, 0.20162};
  *e = 0.0;
  *de = 0.0;
  if (zi <= 0 || zj <= 0 || r <= 0.0 || r >= outer) return;
  const double ainv = 2.134563 * (pow((double)zi, 0.23) + pow((double)zj, 0.23));
  double phi = 0.0, dphi = 0.0;
  for (int k = 0; k < 4; ++k) {
    const double t = a[k] * exp(-b[k] * ainv * r);
    phi += t;
    dphi -= b[k] * ainv * t;
  }
  const double q = 14.3996454784255 * (double)zi * (double)zj;
  double sw = 1.0, dsw = 0.0;
  if (r > inner) {
    const double w = 3.14159265358979323846 / (outer - inner);
    sw = 0.5 * (1.0 + cos(w * (r - inner)));
    dsw = -0.5 * w * sin(w * (r - inner));
  }
  const double u = q * phi / r, du = q * (dphi / r - phi / (r * r));
  *e = sw * u;
  *de = sw * du + dsw * u;
}

/* Per-species atomic numbers and per-pair [S*S] cutoffs of a ZBL model. */
int oracle_set_zbl(oracle_model* m, const int* z, const double* inner, const double* outer)
{
  if (m->S > 16) return -1;
  for (int t = 0; t < m->S; ++t) m->zbl_z[t] = z[t];
  for (int k = 0; k < m->S * m->S; ++k) {
    m->zbl_inner[k] = inner[k];
    m->zbl_outer[k] = outer[k];
  }
  m->zbl_ready = 1;
  return 0;
}

int oracle_zbl_enabled(const oracle_model* m) { return m->zbl; }

/* Basic moments of atom i with the radial coefficients of `mr` (the model, or its copy with
 * the gate radial coefficients), each edge's radial channel mu weighted by mult[j][mu] when
 * mult is given (src/force/sus2_sh.cu:3486-3550, gated form :4246-4310). */
static void or_basics(const oracle_model* mr, const or_nlist* nl, const int* type, int i,
                      const double* mult, double* M)
{
  const int L = mr->L, Mr = mr->Mr;
  double R[64], Y[OR_MAX_COMP];
  for (int b = 0; b < mr->B; ++b) M[b] = 0.0;
  for (int e = nl->start[i]; e < nl->start[i + 1]; ++e) {
    const int j = nl->j[e];
    const double dx = nl->d[3 * e], dy = nl->d[3 * e + 1], dz = nl->d[3 * e + 2];
    const double r = sqrt(dx * dx + dy * dy + dz * dz);
    oracle_radial(mr, type[i], type[j], r, R, NULL);
    if (mult)
      for (int mu = 0; mu < Mr; ++mu) R[mu] *= mult[(size_t)j * Mr + mu];
    oracle_ylm(L, dx, dy, dz, Y, NULL);
    for (int b = 0; b < mr->B; ++b) {
      const int mu = mr->basic_k[b] * (L + 1) + mr->basic_l[b];
      const int y = mr->basic_l[b] * mr->basic_l[b] + mr->basic_m[b] + mr->basic_l[b];
      M[b] += R[mu] * Y[y];
    }
  }
}

/* Forces and virial of the edges of atom i for basic adjoints G (src/force/sus2_sh.cu:
 * 5996-6035, 6347-6446): g_ij = sum_b G_b m_j,mu (R' u Y + R grad Y); F_i += g, F_j -= g,
 * W_i -= r_ij (x) g.  With mult (gated main moments) the gate adjoint
 * adj[j][mu] += dmult[j][mu] sum_b G_b R_mu Y_y is accumulated too (:6077-6123). */
static void or_edge_forces(const oracle_model* mr, const or_nlist* nl, const int* type, int i,
                           const double* G, const double* mult, const double* dmult,
                           double* adj, double* f, int N, double* vir)
{
  const int L = mr->L, Mr = mr->Mr;
  double R[64], dR[64], Y[OR_MAX_COMP], dY[3 * OR_MAX_COMP];
  double fs[3] = {0, 0, 0};
  for (int e = nl->start[i]; e < nl->start[i + 1]; ++e) {
    const int j = nl->j[e];
    const double dx = nl->d[3 * e], dy = nl->d[3 * e + 1], dz = nl->d[3 * e + 2];
    const double r = sqrt(dx * dx + dy * dy + dz * dz);
    const double u[3] = {dx / r, dy / r, dz / r};
    oracle_radial(mr, type[i], type[j], r, R, dR);
    oracle_ylm(L, dx, dy, dz, Y, dY);
    double g[3] = {0, 0, 0};
    for (int b = 0; b < mr->B; ++b) {
      const int mu = mr->basic_k[b] * (L + 1) + mr->basic_l[b];
      const int y = mr->basic_l[b] * mr->basic_l[b] + mr->basic_m[b] + mr->basic_l[b];
      const double w = mult ? mult[(size_t)j * Mr + mu] : 1.0;
      for (int c = 0; c < 3; ++c) g[c] += G[b] * w * (dR[mu] * u[c] * Y[y] + R[mu] * dY[3 * y + c]);
      if (adj) adj[(size_t)j * Mr + mu] += G[b] * R[mu] * Y[y] * dmult[(size_t)j * Mr + mu];
    }
    for (int c = 0; c < 3; ++c) {
      fs[c] += g[c];
      f[(size_t)c * N + j] -= g[c];
    }
    vir[0] -= dx * g[0];
    vir[1] -= dy * g[1];
    vir[2] -= dz * g[2];
    vir[3] -= dx * g[1];
    vir[4] -= dx * g[2];
    vir[5] -= dy * g[2];
    vir[6] -= dy * g[0];
    vir[7] -= dz * g[0];
    vir[8] -= dz * g[1];
  }
  for (int c = 0; c < 3; ++c) f[(size_t)c * N + i] += fs[c];
}

/* Two-layer mu gate evaluation (FP64), in the reference's per-step order
 * (src/force/sus2_sh.cu:9821-10256):
 *   1. gate moments B^g_a (gate radial coefficients), gate graph forward, gate scalars s_q;
 *      h_{a,mu} = sum_q weff[mu][q] s_q(a); mult = 1 + A tanh(a_ta h), dmult = A a (1 - tanh^2)
 *      (:4565-4664)
 *   2. main moments B_i = sum_j mult_j,mu R_mu Y, main graph, E_i, reverse -> G_i (sigma-scaled)
 *   3. forces / virial of the main moments and the gate adjoints adj_{j,mu}
 *   4. seeds dE/ds_q(j) = sum_mu adj_{j,mu} weff[mu][q] (:4666-4790), reverse through the
 *      gate graph -> gate basic adjoints, gate-layer forces / virial (:6695-6808). */
static int or_compute_gate(const oracle_model* m, int N, const int* type, const or_nlist* nl,
                           double* energy, double* force, double* virial)
{
  const int A = m->A, B = m->B, Mr = m->Mr, Qg = m->gate_Qg;
  oracle_model mg = *m; /* same model with the gate radial coefficients */
  mg.radial = m->gate_radial;
  int nthreads = 1;
#ifdef _OPENMP
  nthreads = omp_get_max_threads();
#endif
  double* mult = (double*)calloc((size_t)N * Mr, sizeof(double));
  double* dmult = (double*)calloc((size_t)N * Mr, sizeof(double));
  double* fbuf = (double*)calloc((size_t)nthreads * 3 * N, sizeof(double));
  double* abuf = (double*)calloc((size_t)nthreads * N * Mr, sizeof(double));
  double* adj = (double*)calloc((size_t)N * Mr, sizeof(double));
  if (!mult || !dmult || !fbuf || !abuf || !adj) {
    free(mult), free(dmult), free(fbuf), free(abuf), free(adj);
    return -1;
  }
#pragma omp parallel
  {
    int tid = 0;
#ifdef _OPENMP
    tid = omp_get_thread_num();
#endif
    double* M = (double*)calloc((size_t)A, sizeof(double));
    double* G = (double*)calloc((size_t)A, sizeof(double));
    double* f = fbuf + (size_t)tid * 3 * N;
    double* ad = abuf + (size_t)tid * N * Mr;
    /* 1. gate values */
#pragma omp for schedule(dynamic, 32)
    for (int a = 0; a < N; ++a) {
      for (int k = B; k < A; ++k) M[k] = 0.0;
      or_basics(&mg, nl, type, a, NULL, M);
      or_graph_forward(m, M);
      for (int mu = 0; mu < Mr; ++mu) {
        double h = 0.0;
        for (int q = 0; q < Qg; ++q) h += m->gate_weff[mu * Qg + q] * M[m->gate_node[q]];
        const double ac = m->gate_add[type[a]];
        const double t = tanh(ac * h);
        mult[(size_t)a * Mr + mu] = 1.0 + m->gate_A * t;
        dmult[(size_t)a * Mr + mu] = m->gate_A * ac * (1.0 - t * t);
      }
    }
    /* 2-3. gated main moments, graph, energy, forces and gate adjoints */
#pragma omp for schedule(dynamic, 32)
    for (int i = 0; i < N; ++i) {
      const int ti = type[i];
      for (int k = B; k < A; ++k) M[k] = 0.0;
      or_basics(m, nl, type, i, mult, M);
      const double sigma = m->species[ti];
      energy[i] = m->shift[ti] + sigma + sigma * or_graph_forward(m, M);
      or_graph_reverse(m, sigma, M, G);
      double vir[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
      or_edge_forces(m, nl, type, i, G, mult, dmult, ad, f, N, vir);
      for (int c = 0; c < 9; ++c) virial[(size_t)c * N + i] = vir[c];
    }
#pragma omp for schedule(static)
    for (size_t k = 0; k < (size_t)N * Mr; ++k) {
      double v = 0.0;
      for (int t = 0; t < nthreads; ++t) v += abuf[(size_t)t * N * Mr + k];
      adj[k] = v;
    }
    /* 4. reverse through the gate graph from the per-atom seeds, gate-layer forces */
#pragma omp for schedule(dynamic, 32)
    for (int a = 0; a < N; ++a) {
      for (int k = B; k < A; ++k) M[k] = 0.0;
      or_basics(&mg, nl, type, a, NULL, M);
      or_graph_forward(m, M);
      for (int k = 0; k < A; ++k) G[k] = 0.0;
      for (int q = 0; q < Qg; ++q) {
        double seed = 0.0;
        for (int mu = 0; mu < Mr; ++mu) seed += adj[(size_t)a * Mr + mu] * m->gate_weff[mu * Qg + q];
        G[m->gate_node[q]] += seed;
      }
      for (int p = m->T - 1; p >= 0; --p) {
        const double g = G[m->p_target[p]] * m->p_coeff[p];
        G[m->p_left[p]] += g * M[m->p_right[p]];
        G[m->p_right[p]] += g * M[m->p_left[p]];
      }
      double vir[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
      or_edge_forces(&mg, nl, type, a, G, NULL, NULL, NULL, f, N, vir);
      for (int c = 0; c < 9; ++c) virial[(size_t)c * N + a] += vir[c];
    }
    free(M);
    free(G);
  }
  for (size_t k = 0; k < (size_t)3 * N; ++k) {
    double v = 0.0;
    for (int t = 0; t < nthreads; ++t) v += fbuf[(size_t)t * 3 * N + k];
    force[k] = v;
  }
  free(mult), free(dmult), free(fbuf), free(abuf), free(adj);
  return 0;
}

int oracle_compute(const oracle_model* m, int N, const int* type, const double* pos,
                   const double* h, const int* pbc, int mode, double* energy, double* force,
                   double* virial, double* dparam)
{
  if (m->zbl && !m->zbl_ready) return -1;
  or_init_ylm();
  or_nlist nl;
  if (mode & 1)
    or_neighbors_brute(N, pos, h, pbc, m->rc, &nl);
  else
    or_neighbors_cells(N, pos, h, pbc, m->rc, &nl);
  if (m->gate) {
    /* gate models: energy, forces and virial only (no dE/dtheta, no ZBL) */
    const int rc = (dparam || m->zbl) ? -2 : or_compute_gate(m, N, type, &nl, energy, force, virial);
    or_free_nlist(&nl);
    return rc;
  }

  const int L = m->L, B = m->B, A = m->A, Mr = m->Mr, rb = m->rb, S = m->S;
  int nthreads = 1;
#ifdef _OPENMP
  nthreads = omp_get_max_threads();
#endif
  const int P = oracle_param_count(m);
  double* fbuf = (double*)calloc((size_t)nthreads * 3 * N, sizeof(double));
  double* pbuf = dparam ? (double*)calloc((size_t)nthreads * P, sizeof(double)) : NULL;
  const int off_shift = 0, off_species = S, off_w = 2 * S, off_rad = 2 * S + m->Q,
            off_type = off_rad + Mr * rb, off_scal = off_type + S;

#pragma omp parallel
  {
    int tid = 0;
#ifdef _OPENMP
    tid = omp_get_thread_num();
#endif
    double* f = fbuf + (size_t)tid * 3 * N;
    double* pg = pbuf ? pbuf + (size_t)tid * P : NULL;
    double* M = (double*)malloc((size_t)A * sizeof(double));
    double* G = (double*)malloc((size_t)A * sizeof(double));
    double R[64], dR[64], Y[OR_MAX_COMP], dY[3 * OR_MAX_COMP];
    double Tn[64 * 16], dTn[64 * 16], xi[64], dxa[64], dxs[64];
#pragma omp for schedule(dynamic, 32)
    for (int i = 0; i < N; ++i) {
      const int ti = type[i];
      for (int a = 0; a < A; ++a) M[a] = 0.0;
      /* basic moments */
      for (int e = nl.start[i]; e < nl.start[i + 1]; ++e) {
        const double dx = nl.d[3 * e], dy = nl.d[3 * e + 1], dz = nl.d[3 * e + 2];
        const double r = sqrt(dx * dx + dy * dy + dz * dz);
        oracle_radial(m, ti, type[nl.j[e]], r, R, NULL);
        oracle_ylm(L, dx, dy, dz, Y, NULL);
        for (int b = 0; b < B; ++b) {
          const int mu = m->basic_k[b] * (L + 1) + m->basic_l[b];
          const int y = m->basic_l[b] * m->basic_l[b] + m->basic_m[b] + m->basic_l[b];
          M[b] += R[mu] * Y[y];
        }
      }
      /* flat product graph */
      const double sigma = m->species[ti];
      const double sum = or_graph_forward(m, M);
      energy[i] = m->shift[ti] + sigma + sigma * sum;
      if (pg) {
        pg[off_shift + ti] += 1.0;
        pg[off_species + ti] += 1.0 + sum;
        for (int s = 0; s < m->Q; ++s) pg[off_w + s] += sigma * M[m->mapping[s]];
      }
      /* reverse */
      or_graph_reverse(m, sigma, M, G);
      /* forces and virial */
      double vir[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
      double fs[3] = {0, 0, 0};
      for (int e = nl.start[i]; e < nl.start[i + 1]; ++e) {
        const int j = nl.j[e];
        const int tj = type[j];
        const double dx = nl.d[3 * e], dy = nl.d[3 * e + 1], dz = nl.d[3 * e + 2];
        const double r = sqrt(dx * dx + dy * dy + dz * dz);
        const double u[3] = {dx / r, dy / r, dz / r};
        or_radial_full(m, ti, tj, r, R, dR, pg ? Tn : NULL, pg ? dTn : NULL, xi, dxa, dxs);
        oracle_ylm(L, dx, dy, dz, Y, dY);
        double g[3] = {0, 0, 0};
        for (int b = 0; b < B; ++b) {
          const double adj = G[b];
          const int mu = m->basic_k[b] * (L + 1) + m->basic_l[b];
          const int y = m->basic_l[b] * m->basic_l[b] + m->basic_m[b] + m->basic_l[b];
          for (int c = 0; c < 3; ++c) g[c] += adj * (dR[mu] * u[c] * Y[y] + R[mu] * dY[3 * y + c]);
          if (pg) {
            /* dB_b/dtheta for radial coefficient c_{mu n} = scaling w_ti w_tj rad[mu][n] */
            const double tsc = m->scaling * m->type_coeffs[ti] * m->type_coeffs[tj];
            const int so = 2 * mu * S * S + S * ti + tj;
            double dRda = 0.0, dRds = 0.0; /* d R_mu / d a, d s (through xi) */
            for (int n = 0; n < rb; ++n) {
              const double cr = m->radial[mu * rb + n];
              pg[off_rad + mu * rb + n] += adj * tsc * Tn[mu * rb + n] * Y[y];
              /* type coefficients enter via tsc = scaling w_ti w_tj */
              const double dval = adj * m->scaling * cr * Tn[mu * rb + n] * Y[y];
              pg[off_type + ti] += dval * m->type_coeffs[tj];
              pg[off_type + tj] += dval * m->type_coeffs[ti];
              dRda += cr * tsc * dTn[mu * rb + n] * dxa[mu];
              dRds += cr * tsc * dTn[mu * rb + n] * dxs[mu];
            }
            pg[off_scal + so] += adj * dRda * Y[y];
            pg[off_scal + so + S * S] += adj * dRds * Y[y];
          }
        }
        fs[0] += g[0];
        fs[1] += g[1];
        fs[2] += g[2];
        f[j] -= g[0];
        f[N + j] -= g[1];
        f[2 * N + j] -= g[2];
        vir[0] -= dx * g[0];
        vir[1] -= dy * g[1];
        vir[2] -= dz * g[2];
        vir[3] -= dx * g[1];
        vir[4] -= dx * g[2];
        vir[5] -= dy * g[2];
        vir[6] -= dy * g[0];
        vir[7] -= dz * g[0];
        vir[8] -= dz * g[1];
      }
      /* ZBL: E_i += phi / 2 per edge inside the pair cutoff, g = (phi' / 2) d / r */
      for (int e = nl.start[i]; m->zbl && e < nl.start[i + 1]; ++e) {
        const int j = nl.j[e];
        const int pair = ti * S + type[j];
        const double dx = nl.d[3 * e], dy = nl.d[3 * e + 1], dz = nl.d[3 * e + 2];
        const double r = sqrt(dx * dx + dy * dy + dz * dz);
        double ez, dez;
        oracle_zbl_pair(m->zbl_z[ti], m->zbl_z[type[j]], r, m->zbl_inner[pair], m->zbl_outer[pair],
                        &ez, &dez);
        if (ez == 0.0 && dez == 0.0) continue;
        energy[i] += 0.5 * ez;
        const double g[3] = {0.5 * dez * dx / r, 0.5 * dez * dy / r, 0.5 * dez * dz / r};
        fs[0] += g[0];
        fs[1] += g[1];
        fs[2] += g[2];
        f[j] -= g[0];
        f[N + j] -= g[1];
        f[2 * N + j] -= g[2];
        vir[0] -= dx * g[0];
        vir[1] -= dy * g[1];
        vir[2] -= dz * g[2];
        vir[3] -= dx * g[1];
        vir[4] -= dx * g[2];
        vir[5] -= dy * g[2];
        vir[6] -= dy * g[0];
        vir[7] -= dz * g[0];
        vir[8] -= dz * g[1];
      }
      f[i] += fs[0];
      f[N + i] += fs[1];
      f[2 * N + i] += fs[2];
      for (int c = 0; c < 9; ++c) virial[c * N + i] = vir[c];
    }
    free(M);
    free(G);
  }
  for (int k = 0; k < 3 * N; ++k) {
    double s = 0.0;
    for (int t = 0; t < nthreads; ++t) s += fbuf[(size_t)t * 3 * N + k];
    force[k] = s;
  }
  if (dparam) {
    for (int k = 0; k < P; ++k) {
      double s = 0.0;
      for (int t = 0; t < nthreads; ++t) s += pbuf[(size_t)t * P + k];
      dparam[k] = s;
    }
  }
  free(fbuf);
  free(pbuf);
  or_free_nlist(&nl);
  return 0;
}

int oracle_num_threads(void)
{
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void oracle_set_num_threads(int n)
{
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}
