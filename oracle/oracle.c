/*
 * ORACLE — CPU restatement of Grico.jl's matrix-free cell algorithm, FP64.
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the
 * checker or the timed CPU baseline; the product path never calls it.
 *
 * Every routine follows the reference line by line in structure:
 *   gather           src/assembly.jl:1751-1771 (_gather_local_coefficients!,
 *                    single-term fast path src/common.jl:334-370)
 *   interpolate      src/integration.jl:1140-1155 (_tensor_interpolate_with_axis!:
 *                    axes 1..D in order, derivative table on the derivative axis),
 *                    contraction src/integration.jl:995-1016 (muladd accumulation)
 *   gradient tables  src/integration.jl:2707-2708 (axis_gradients = J^-1 * psi')
 *   project          src/integration.jl:1266-1287 / 1378-1399 (transposed passes,
 *                    one per channel, accumulated into the mode box)
 *   scatter          src/assembly.jl:1788-1803 (skip zero values, skip fixed rows)
 *   constraint rows  src/assembly.jl:1893-1920 (fixed rows = identity)
 *   basis path       src/accumulators.jl:1022-1054 (full mode x point tables)
 *   diagonal         src/accumulators.jl:1056-1083 + src/assembly.jl:1811-1825
 *   CG               src/solve.jl:376-431
 *
 * Per-point actions: the reference calls the user's cell_accumulate at each
 * point. The two-phase entry points (oracle_interpolate / oracle_project)
 * keep that structure: the caller evaluates the callback on the trial
 * channels these routines produce. oracle_apply_lowered takes the action as a
 * per-point (D+1)x(D+1) matrix instead (linear accumulators) and runs the
 * whole pipeline in C; it is the timed CPU baseline.
 *
 * Data layout of one batch ("bucket") of cells with equal degree tuple and
 * quadrature shape: mode boxes and point sets are ordered axis 1 fastest;
 * tables[a] is N_a x Q_a row-major [mode][point]; box entries map to global
 * dofs through a CSR (box_term_offsets over cells*Ntot entries).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct oracle_bucket {
  int dim;
  int N[3];
  int Q[3];
  int64_t cells;
  const double* values[3];       /* N_a x Q_a */
  const double* derivatives[3];  /* N_a x Q_a, reference coordinates */
  const double* inverse_jacobian; /* cells x dim */
  const double* weights;          /* cells x Qtot (W_q = w_q detJ) */
  const int64_t* term_offsets;    /* cells*Ntot + 1 */
  const int64_t* term_indices;
  const double* term_coefficients;
} oracle_bucket;

static int ntot(const oracle_bucket* b) {
  int n = 1;
  for (int a = 0; a < b->dim; ++a) n *= b->N[a];
  return n;
}

static int qtot(const oracle_bucket* b) {
  int n = 1;
  for (int a = 0; a < b->dim; ++a) n *= b->Q[a];
  return n;
}

/* One contraction along `axis` (src/integration.jl:995-1016):
 * out[..q..] = sum_i T[i][q] in[..i..] (forward) or
 * out[..i..] = sum_q T[i][q] in[..q..] (transpose, 1065-1087). */
static void contract(double* out, const double* in, const int* in_shape, int D, int axis, int out_ext,
                     const double* T, int tq, int transpose) {
  int stride = 1, outer = 1;
  for (int a = 0; a < D; ++a) {
    if (a < axis) stride *= in_shape[a];
    if (a > axis) outer *= in_shape[a];
  }
  int in_ext = in_shape[axis];
  for (int hi = 0; hi < outer; ++hi)
    for (int j = 0; j < out_ext; ++j)
      for (int lo = 0; lo < stride; ++lo) {
        double v = 0.0;
        const double* src = in + lo + (int64_t)hi * stride * in_ext;
        for (int k = 0; k < in_ext; ++k) {
          double t = transpose ? T[j * tq + k] : T[k * tq + j];
          v = fma(t, src[k * stride], v);
        }
        out[lo + stride * (j + out_ext * hi)] = v;
      }
}

typedef struct scratch {
  double* a;
  double* b;
  double* tables; /* scaled gradient tables per axis */
} scratch;

static int box_max(const oracle_bucket* b) {
  int m = 1;
  for (int a = 0; a < b->dim; ++a) m *= (b->N[a] > b->Q[a] ? b->N[a] : b->Q[a]);
  return m;
}

static void scratch_init(scratch* s, const oracle_bucket* b) {
  int m = box_max(b);
  s->a = (double*)malloc(sizeof(double) * (size_t)m);
  s->b = (double*)malloc(sizeof(double) * (size_t)m);
  int t = 0;
  for (int a = 0; a < b->dim; ++a) t += b->N[a] * b->Q[a];
  s->tables = (double*)malloc(sizeof(double) * (size_t)(t > 0 ? t : 1));
}

static void scratch_free(scratch* s) {
  free(s->a);
  free(s->b);
  free(s->tables);
}

/* J^-1-scaled gradient tables of one cell (src/integration.jl:2707-2708). */
static void scaled_gradients(const oracle_bucket* b, int64_t cell, double* out, const double** per_axis) {
  int off = 0;
  for (int a = 0; a < b->dim; ++a) {
    double j = b->inverse_jacobian[cell * b->dim + a];
    int n = b->N[a] * b->Q[a];
    for (int k = 0; k < n; ++k) out[off + k] = j * b->derivatives[a][k];
    per_axis[a] = out + off;
    off += n;
  }
}

/* Interpolate with the derivative table on `daxis` (-1: value) into dst. */
static void interpolate_axis(const oracle_bucket* b, const double* coef, double* dst, int daxis,
                             const double** grads, scratch* s) {
  int shape[3] = {1, 1, 1};
  for (int a = 0; a < b->dim; ++a) shape[a] = b->N[a];
  const double* cur = coef;
  double* bufs[2] = {s->a, s->b};
  for (int a = 0; a < b->dim; ++a) {
    double* nxt = (a == b->dim - 1) ? dst : bufs[a & 1];
    const double* T = (a == daxis) ? grads[a] : b->values[a];
    contract(nxt, cur, shape, b->dim, a, b->Q[a], T, b->Q[a], 0);
    shape[a] = b->Q[a];
    cur = nxt;
  }
}

/* result[mode] += projection of point data with the derivative table on daxis. */
static void project_axis(const oracle_bucket* b, const double* f, double* result, int daxis,
                         const double** grads, scratch* s) {
  int shape[3] = {1, 1, 1};
  for (int a = 0; a < b->dim; ++a) shape[a] = b->Q[a];
  const double* cur = f;
  double* bufs[2] = {s->a, s->b};
  for (int a = 0; a < b->dim; ++a) {
    double* nxt = bufs[a & 1];
    const double* T = (a == daxis) ? grads[a] : b->values[a];
    contract(nxt, cur, shape, b->dim, a, b->N[a], T, b->Q[a], 1);
    shape[a] = b->N[a];
    cur = nxt;
  }
  int n = ntot(b);
  for (int i = 0; i < n; ++i) result[i] += cur[i];
}

static void gather(const oracle_bucket* b, int64_t cell, const double* x, double* coef) {
  int n = ntot(b);
  for (int i = 0; i < n; ++i) {
    int64_t e = cell * n + i;
    int64_t t0 = b->term_offsets[e], t1 = b->term_offsets[e + 1];
    double v = 0.0;
    if (t1 - t0 == 1) {
      v = b->term_coefficients[t0] * x[b->term_indices[t0]];
    } else {
      for (int64_t t = t0; t < t1; ++t) v += b->term_coefficients[t] * x[b->term_indices[t]];
    }
    coef[i] = v;
  }
}

static void scatter(const oracle_bucket* b, int64_t cell, const double* r, const uint8_t* fixed, double* y,
                    int squared) {
  int n = ntot(b);
  for (int i = 0; i < n; ++i) {
    double v = r[i];
    if (v == 0.0) continue;
    int64_t e = cell * n + i;
    for (int64_t t = b->term_offsets[e]; t < b->term_offsets[e + 1]; ++t) {
      int64_t row = b->term_indices[t];
      if (fixed && fixed[row]) continue;
      double c = b->term_coefficients[t];
      y[row] += (squared ? c * c : c) * v;
    }
  }
}

/* Trial channels at every point: values (cells x Q) and physical gradients
 * (cells x Q x D) of the field with full-layout coefficients x. */
void oracle_interpolate(const oracle_bucket* b, const double* x, double* values, double* gradients) {
  int n = ntot(b), q = qtot(b), D = b->dim;
  scratch s;
  scratch_init(&s, b);
  double* coef = (double*)malloc(sizeof(double) * (size_t)n);
  double* tmp = (double*)malloc(sizeof(double) * (size_t)q);
  const double* grads[3];
  for (int64_t c = 0; c < b->cells; ++c) {
    gather(b, c, x, coef);
    scaled_gradients(b, c, s.tables, grads);
    interpolate_axis(b, coef, values + c * q, -1, grads, &s);
    for (int a = 0; a < D; ++a) {
      interpolate_axis(b, coef, tmp, a, grads, &s);
      for (int k = 0; k < q; ++k) gradients[(c * q + k) * D + a] = tmp[k];
    }
  }
  free(tmp);
  free(coef);
  scratch_free(&s);
}

/* Project weighted test channels (a0 * W, a_grad * W) and scatter into y,
 * skipping rows with fixed[row] != 0 (fixed may be NULL). */
void oracle_project(const oracle_bucket* b, const double* a0w, const double* agw, const uint8_t* fixed, double* y) {
  int n = ntot(b), q = qtot(b), D = b->dim;
  scratch s;
  scratch_init(&s, b);
  double* r = (double*)malloc(sizeof(double) * (size_t)n);
  double* tmp = (double*)malloc(sizeof(double) * (size_t)q);
  const double* grads[3];
  for (int64_t c = 0; c < b->cells; ++c) {
    memset(r, 0, sizeof(double) * (size_t)n);
    scaled_gradients(b, c, s.tables, grads);
    project_axis(b, a0w + c * q, r, -1, grads, &s);
    for (int a = 0; a < D; ++a) {
      for (int k = 0; k < q; ++k) tmp[k] = agw[(c * q + k) * D + a];
      project_axis(b, tmp, r, a, grads, &s);
    }
    scatter(b, c, r, fixed, y, 0);
  }
  free(tmp);
  free(r);
  scratch_free(&s);
}

/* Fixed rows are identities (src/assembly.jl:1897-1899). */
void oracle_fixed_rows(const int64_t* fixed, int64_t count, const double* x, double* y) {
  for (int64_t i = 0; i < count; ++i) y[fixed[i]] = x[fixed[i]];
}

/* Whole pipeline with a lowered linear action: per point
 * (a0, a_grad) = M_q (u, grad u), M_q row-major (D+1)x(D+1) unweighted;
 * weighted by W_q as the reference does after the callback. Cells are
 * processed in `colors` batches (DOF-disjoint within a batch) so the batch
 * loop can run threaded like the reference's colored traversal
 * (src/assembly.jl:1350-1374). */
void oracle_apply_lowered(const oracle_bucket* b, const double* M, const double* x, const uint8_t* fixed, double* y,
                          const int64_t* color_offsets, const int64_t* color_cells, int colors, int threads) {
  int n = ntot(b), q = qtot(b), D = b->dim, C = D + 1;
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#else
  (void)threads;
#endif
  for (int col = 0; col < colors; ++col) {
    int64_t lo = color_offsets[col], hi = color_offsets[col + 1];
#pragma omp parallel
    {
      scratch s;
      scratch_init(&s, b);
      double* coef = (double*)malloc(sizeof(double) * (size_t)n);
      double* r = (double*)malloc(sizeof(double) * (size_t)n);
      double* ch = (double*)malloc(sizeof(double) * (size_t)(q * C));
      double* tmp = (double*)malloc(sizeof(double) * (size_t)q);
      const double* grads[3];
#pragma omp for schedule(static)
      for (int64_t k = lo; k < hi; ++k) {
        int64_t c = color_cells[k];
        gather(b, c, x, coef);
        scaled_gradients(b, c, s.tables, grads);
        interpolate_axis(b, coef, ch, -1, grads, &s);
        for (int a = 0; a < D; ++a) interpolate_axis(b, coef, ch + (a + 1) * q, a, grads, &s);
        for (int p = 0; p < q; ++p) {
          const double* Mq = M + ((c * q) + p) * C * C;
          double w = b->weights[c * q + p];
          double in[4], out[4];
          for (int i = 0; i < C; ++i) in[i] = ch[i * q + p];
          for (int i = 0; i < C; ++i) {
            double v = 0.0;
            for (int j = 0; j < C; ++j) v += Mq[i * C + j] * in[j];
            out[i] = w * v;
          }
          for (int i = 0; i < C; ++i) ch[i * q + p] = out[i];
        }
        memset(r, 0, sizeof(double) * (size_t)n);
        project_axis(b, ch, r, -1, grads, &s);
        for (int a = 0; a < D; ++a) {
          memcpy(tmp, ch + (a + 1) * q, sizeof(double) * (size_t)q);
          project_axis(b, tmp, r, a, grads, &s);
        }
        scatter(b, c, r, fixed, y, 0);
      }
      free(tmp);
      free(ch);
      free(r);
      free(coef);
      scratch_free(&s);
    }
  }
}

/* Brute-force basis path with full tables (src/accumulators.jl:1022-1054):
 * phi[i][q] = prod_a psi_{i_a}(xi_q), grad phi via J^-1-scaled derivatives. */
static void basis_tables_cell(const oracle_bucket* b, int64_t cell, double* phi, double* dphi) {
  int n = ntot(b), q = qtot(b), D = b->dim;
  for (int i = 0; i < n; ++i) {
    int mi[3] = {0, 0, 0}, rem = i;
    for (int a = 0; a < D; ++a) {
      mi[a] = rem % b->N[a];
      rem /= b->N[a];
    }
    for (int p = 0; p < q; ++p) {
      int qi[3] = {0, 0, 0}, rp = p;
      for (int a = 0; a < D; ++a) {
        qi[a] = rp % b->Q[a];
        rp /= b->Q[a];
      }
      double v = 1.0;
      for (int a = 0; a < D; ++a) v *= b->values[a][mi[a] * b->Q[a] + qi[a]];
      phi[i * q + p] = v;
      for (int d = 0; d < D; ++d) {
        double g = b->inverse_jacobian[cell * D + d] * b->derivatives[d][mi[d] * b->Q[d] + qi[d]];
        for (int a = 0; a < D; ++a)
          if (a != d) g *= b->values[a][mi[a] * b->Q[a] + qi[a]];
        dphi[(d * n + i) * q + p] = g;
      }
    }
  }
}

void oracle_apply_basis(const oracle_bucket* b, const double* M, const double* x, const uint8_t* fixed, double* y) {
  int n = ntot(b), q = qtot(b), D = b->dim, C = D + 1;
  double* phi = (double*)malloc(sizeof(double) * (size_t)(n * q));
  double* dphi = (double*)malloc(sizeof(double) * (size_t)(D * n * q));
  double* coef = (double*)malloc(sizeof(double) * (size_t)n);
  double* r = (double*)malloc(sizeof(double) * (size_t)n);
  for (int64_t c = 0; c < b->cells; ++c) {
    basis_tables_cell(b, c, phi, dphi);
    gather(b, c, x, coef);
    memset(r, 0, sizeof(double) * (size_t)n);
    for (int p = 0; p < q; ++p) {
      double in[4] = {0, 0, 0, 0}, out[4];
      for (int i = 0; i < n; ++i) {
        in[0] += phi[i * q + p] * coef[i];
        for (int d = 0; d < D; ++d) in[d + 1] += dphi[(d * n + i) * q + p] * coef[i];
      }
      const double* Mq = M + ((c * q) + p) * C * C;
      for (int i = 0; i < C; ++i) {
        double v = 0.0;
        for (int j = 0; j < C; ++j) v += Mq[i * C + j] * in[j];
        out[i] = v;
      }
      double w = b->weights[c * q + p];
      for (int i = 0; i < n; ++i) {
        double t = phi[i * q + p] * out[0];
        for (int d = 0; d < D; ++d) t += dphi[(d * n + i) * q + p] * out[d + 1];
        r[i] += w * t;
      }
    }
    scatter(b, c, r, fixed, y, 0);
  }
  free(r);
  free(coef);
  free(dphi);
  free(phi);
}

/* Local diagonal sum_q W [phi_i a0(phi_i) + grad phi_i . a_grad(phi_i)],
 * scattered with squared coefficients (src/assembly.jl:1811-1825). */
void oracle_diagonal(const oracle_bucket* b, const double* M, double* diag) {
  int n = ntot(b), q = qtot(b), D = b->dim, C = D + 1;
  double* phi = (double*)malloc(sizeof(double) * (size_t)(n * q));
  double* dphi = (double*)malloc(sizeof(double) * (size_t)(D * n * q));
  double* r = (double*)malloc(sizeof(double) * (size_t)n);
  for (int64_t c = 0; c < b->cells; ++c) {
    basis_tables_cell(b, c, phi, dphi);
    for (int i = 0; i < n; ++i) {
      double acc = 0.0;
      for (int p = 0; p < q; ++p) {
        double in[4];
        in[0] = phi[i * q + p];
        for (int d = 0; d < D; ++d) in[d + 1] = dphi[(d * n + i) * q + p];
        const double* Mq = M + ((c * q) + p) * C * C;
        double t = 0.0;
        for (int k = 0; k < C; ++k) {
          double v = 0.0;
          for (int j = 0; j < C; ++j) v += Mq[k * C + j] * in[j];
          t += in[k] * v;
        }
        acc += b->weights[c * q + p] * t;
      }
      r[i] = acc;
    }
    scatter(b, c, r, NULL, diag, 1);
  }
  free(r);
  free(dphi);
  free(phi);
}

/* Greedy DOF-disjoint coloring in cell order (src/plans.jl:368-396): each
 * cell takes the first color none of its dofs has been claimed by. */
int oracle_color_cells(const oracle_bucket* b, int64_t ndofs, int32_t* color_of_cell) {
  int n = ntot(b);
  uint64_t* used = (uint64_t*)calloc((size_t)ndofs, sizeof(uint64_t));
  int max_color = 0;
  for (int64_t c = 0; c < b->cells; ++c) {
    uint64_t mask = 0;
    for (int64_t t = b->term_offsets[c * n]; t < b->term_offsets[(c + 1) * n]; ++t) mask |= used[b->term_indices[t]];
    int col = 0;
    while (col < 64 && (mask >> col & 1u)) ++col;
    if (col == 64) {
      free(used);
      return -1;
    }
    color_of_cell[c] = col;
    if (col + 1 > max_color) max_color = col + 1;
    for (int64_t t = b->term_offsets[c * n]; t < b->term_offsets[(c + 1) * n]; ++t)
      used[b->term_indices[t]] |= (uint64_t)1 << col;
  }
  free(used);
  return max_color;
}

static double dotp(const double* a, const double* b, int64_t n) {
  double s = 0.0;
  for (int64_t i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

typedef void (*oracle_operator)(void* ctx, const double* x, double* y);

/* Preconditioned CG (src/solve.jl:376-431). inv_diag may be NULL. Returns
 * iterations >= 0, -1 on non-convergence, -2 on a non-positive direction. */
int64_t oracle_cg(oracle_operator op, void* ctx, int64_t n, const double* inv_diag, const double* b, double* x,
                  double rtol, double atol, int64_t maxiter) {
  double* r = (double*)malloc(sizeof(double) * (size_t)n);
  double* z = (double*)malloc(sizeof(double) * (size_t)n);
  double* p = (double*)malloc(sizeof(double) * (size_t)n);
  double* q = (double*)malloc(sizeof(double) * (size_t)n);
  int64_t result = -1;
  memset(x, 0, sizeof(double) * (size_t)n);
  memcpy(r, b, sizeof(double) * (size_t)n);
  for (int64_t i = 0; i < n; ++i) z[i] = inv_diag ? inv_diag[i] * r[i] : r[i];
  memcpy(p, z, sizeof(double) * (size_t)n);
  double rz = dotp(r, z, n);
  double bnorm = sqrt(dotp(b, b, n));
  double tol = atol > rtol * bnorm ? atol : rtol * bnorm;
  double rr = dotp(r, r, n);
  if (rr <= tol * tol) {
    result = 0;
    goto done;
  }
  for (int64_t it = 1; it <= maxiter; ++it) {
    op(ctx, p, q);
    double den = dotp(p, q, n);
    if (!(den > 0.0)) {
      result = -2;
      goto done;
    }
    double alpha = rz / den;
    for (int64_t i = 0; i < n; ++i) {
      x[i] += alpha * p[i];
      r[i] -= alpha * q[i];
    }
    rr = dotp(r, r, n);
    if (rr <= tol * tol) {
      result = it;
      goto done;
    }
    for (int64_t i = 0; i < n; ++i) z[i] = inv_diag ? inv_diag[i] * r[i] : r[i];
    double rz2 = dotp(r, z, n);
    double beta = rz2 / rz;
    for (int64_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    rz = rz2;
  }
done:
  free(q);
  free(p);
  free(z);
  free(r);
  return result;
}
