/*
 * ubt_oracle.c - TEST INFRASTRUCTURE ONLY. A plain-C FP64 restatement of the
 * reference's (stevendargaville/UBOLT) transport operator path, used by the
 * tests as the parity checker and by bench.py as the timed CPU baseline. The
 * product (libubolt_amd) never links, loads or calls this file.
 *
 * It keeps the reference's structure rather than the product's: the streaming
 * and removal terms are assembled into an explicit sparse matrix row by row,
 * slot by slot, exactly as the reference's COO pattern and value fills lay
 * them out, the operator mult is that SpMV plus the matrix-free scatter, and
 * the element blocks are read back out of the assembled rows before being
 * inverted by Gauss-Jordan - the same three steps as ShellMatMultApply and
 * ElementBlockInverse. Every function cites the reference lines it restates
 * (paths relative to the reference root).
 *
 * Layouts are the reference's: row = (cell * nb + basis) * A + angle; omega 3
 * per angle; reflect 3 per angle (reflect[3a+d] mirrors cosine d); face tables
 * per slot k = cell * nf + face; face_nb >= 0 neighbour cell, -1 vacuum,
 * -2-axis reflective across the face's dominant axis.
 *
 * OpenMP parallelises over rows / cells only; every reduction is per row, so
 * results are independent of the thread count.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

typedef int64_t i64;

int oracle_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void oracle_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

/* ------------------------------------------------------------------------
 * Structured FD, ghost-flux treatment.
 * Slots per row: one upwind slot per axis in axis order, then the diagonal
 * (src/structured_fd_commonk.hpp:286-397); values |cos_d|/h_d
 * (src/termsk.kokkos.cxx:70-98) plus sigma_t on the diagonal
 * (src/termsk.kokkos.cxx:451-471). A nulled slot has col -1.
 * ------------------------------------------------------------------------ */
int oracle_fd_slots(int dim) { return dim + 1; }

void oracle_fd_assemble(int dim, const int *n, const double *h, int A, const double *omega, const int *reflect,
                        const int *face_type /* [d*2+side], 1 = reflect */, const double *sigt, i64 *col,
                        double *val) {
  const i64 cells = (i64)n[0] * n[1] * (dim == 3 ? n[2] : 1);
  const int S = dim + 1;
#pragma omp parallel for schedule(static)
  for (i64 c = 0; c < cells; c++) {
    int idx[3];
    i64 cc = c;
    idx[0] = (int)(cc % n[0]);
    cc /= n[0];
    idx[1] = (int)(cc % n[1]);
    idx[2] = dim == 3 ? (int)(cc / n[1]) : 0;
    const i64 stride[3] = {1, n[0], (i64)n[0] * n[1]};
    for (int a = 0; a < A; a++) {
      const i64 r = c * A + a;
      double cd[3] = {0.0, 0.0, 0.0}, diag = 0.0;
      for (int d = 0; d < dim; d++) cd[d] = fabs(omega[3 * a + d]) / h[d];
      diag = cd[0];
      for (int d = 1; d < dim; d++) diag += cd[d];
      for (int d = 0; d < dim; d++) {
        const double cs = omega[3 * a + d];
        i64 cl = -1;
        if (cs != 0.0) {
          const int up = cs > 0.0 ? idx[d] - 1 : idx[d] + 1;
          if (up >= 0 && up < n[d])
            cl = (c + (cs > 0.0 ? -stride[d] : stride[d])) * A + a;
          else if (face_type[2 * d + (cs > 0.0 ? 0 : 1)] == 1)
            cl = c * A + reflect[3 * a + d];
        }
        col[r * S + d] = cl;
        val[r * S + d] = cl >= 0 ? -cd[d] : 0.0;
      }
      col[r * S + dim] = r;
      val[r * S + dim] = diag + sigt[c];
    }
  }
}

/* Ghost inflow of the structured box: per row, for every axis whose upwind
 * neighbour is outside the box through a vacuum face, |cos_d|/h_d times that
 * face's inflow/W if the cell centre is inside its window, summed in axis
 * order (src/structured_fd_commonk.hpp:330-343). face_bc: per axis side 7
 * doubles: type, inflow/W, n_window_pairs, window[4]. */
void oracle_fd_ghost_inflow(int dim, const int *n, const double *h, int A, const double *omega,
                            const double *face_bc, double *b) {
  const i64 cells = (i64)n[0] * n[1] * (dim == 3 ? n[2] : 1);
#pragma omp parallel for schedule(static)
  for (i64 c = 0; c < cells; c++) {
    int idx[3];
    i64 cc = c;
    idx[0] = (int)(cc % n[0]);
    cc /= n[0];
    idx[1] = (int)(cc % n[1]);
    idx[2] = dim == 3 ? (int)(cc / n[1]) : 0;
    double centre[3];
    for (int d = 0; d < dim; d++) centre[d] = (idx[d] + 0.5) * h[d];
    for (int a = 0; a < A; a++) {
      double v = 0.0;
      for (int d = 0; d < dim; d++) {
        const double cs = omega[3 * a + d];
        if (cs == 0.0) continue;
        const int up = cs > 0.0 ? idx[d] - 1 : idx[d] + 1;
        if (up >= 0 && up < n[d]) continue;
        const double *f = face_bc + 7 * (2 * d + (cs > 0.0 ? 0 : 1));
        if ((int)f[0] != 0) continue; /* reflective: a coupling, not an inflow */
        int inside = 1, p = 0;
        for (int e = 0; e < dim; e++) {
          if (e == d) continue;
          if (p < (int)f[2]) inside = inside && centre[e] >= f[3 + 2 * p] && centre[e] <= f[4 + 2 * p];
          p++;
        }
        if (inside) v += fabs(cs) / h[d] * f[1];
      }
      b[c * A + a] += v;
    }
  }
}

/* ------------------------------------------------------------------------
 * Unstructured DG0 / DG1, ghost-flux treatment.
 * Slots per row: nb per face in cone order (the upwind cell's basis j on an
 * interior inflow face, this cell's basis j at the mirrored angle on a
 * reflective inflow face, nulled otherwise), then this cell's own nb
 * (src/unstructured_dgk.kokkos.cxx:442-520, DG0 :536-580). Values:
 * DG0 src/termsk.kokkos.cxx:197-235, DG1 :328-385, removal :451-471.
 * ------------------------------------------------------------------------ */
int oracle_dg_slots(int nb, int nf) { return (nf + 1) * nb; }

static double face_flux(const double *omega, int a, const double *nA, i64 k) {
  return omega[3 * a] * nA[3 * k] + omega[3 * a + 1] * nA[3 * k + 1] + omega[3 * a + 2] * nA[3 * k + 2];
}

void oracle_dg_assemble(int nb, int nf, i64 cells, int A, const double *omega, const int *reflect,
                        const int *face_nb, const double *face_nA, const double *volume, const double *face_own,
                        const double *face_up, const double *basis_grad, const double *sigt, i64 *col,
                        double *val) {
  const int S = (nf + 1) * nb;
#pragma omp parallel for schedule(static)
  for (i64 c = 0; c < cells; c++) {
    const double iv = 1.0 / volume[c];
    for (int i = 0; i < nb; i++)
      for (int a = 0; a < A; a++) {
        const i64 r = (c * nb + i) * A + a;
        i64 *cl = col + r * S;
        double *vl = val + r * S;
        const int own = nf * nb;
        for (int s = 0; s < S; s++) {
          cl[s] = -1;
          vl[s] = 0.0;
        }
        for (int j = 0; j < nb; j++) cl[own + j] = (c * nb + j) * A + a;
        double outflow = 0.0, diag1 = 0.0;
        for (int f = 0; f < nf; f++) {
          const i64 k = c * nf + f;
          const double s = face_flux(omega, a, face_nA, k);
          if (s > 0.0) {
            if (nb == 1)
              outflow += s;
            else {
              diag1 += s * face_own[(k * nb + i) * nb + i];
              for (int j = 0; j < nb; j++)
                if (j != i) vl[own + j] += s * face_own[(k * nb + i) * nb + j];
            }
            continue;
          }
          const int nbk = face_nb[k];
          if (nbk == -1 || !(s < 0.0)) continue;
          for (int j = 0; j < nb; j++) {
            const i64 colj = nbk >= 0 ? ((i64)nbk * nb + j) * A + a : (c * nb + j) * A + reflect[3 * a + (-2 - nbk)];
            cl[f * nb + j] = colj;
            vl[f * nb + j] = nb == 1 ? s * iv : s * face_up[(k * nb + i) * nb + j];
          }
        }
        if (nb > 1 && i > 0)
          vl[own] -= omega[3 * a] * basis_grad[(c * nb + i) * 3] + omega[3 * a + 1] * basis_grad[(c * nb + i) * 3 + 1] +
                     omega[3 * a + 2] * basis_grad[(c * nb + i) * 3 + 2];
        vl[own + i] += nb == 1 ? outflow * iv : diag1;
        vl[own + i] += sigt[c];
      }
  }
}

/* DG ghost inflow: -(s/V) phi_i(x_f) inflow/W on every vacuum face the angle
 * flows in through (src/unstructured_dgk.kokkos.cxx:490-493, DG0 :560-567).
 * face_inflow is inflow/W on a vacuum face whose centroid is in its window. */
void oracle_dg_ghost_inflow(int nb, int nf, i64 cells, int A, const double *omega, const int *face_nb,
                            const double *face_nA, const double *volume, const double *face_basis_value,
                            const double *face_inflow, double *b) {
#pragma omp parallel for schedule(static)
  for (i64 c = 0; c < cells; c++)
    for (int f = 0; f < nf; f++) {
      const i64 k = c * nf + f;
      if (face_nb[k] != -1 || face_inflow[k] == 0.0) continue;
      for (int a = 0; a < A; a++) {
        const double s = face_flux(omega, a, face_nA, k);
        if (!(s < 0.0)) continue;
        for (int i = 0; i < nb; i++) b[(c * nb + i) * A + a] += -s / volume[c] * face_basis_value[k * nb + i] * face_inflow[k];
      }
    }
}

/* ------------------------------------------------------------------------
 * The shell mult: assembled rows times x, then the matrix-free scatter
 * (src/transport_operatork.kokkos.cxx:49-61; scatter src/termsk.kokkos.cxx:
 * 584-649 with the angular integral src/sn_quadraturek.kokkos.cxx:400-427).
 * ------------------------------------------------------------------------ */
void oracle_rows_mult(i64 rows, int S, const i64 *col, const double *val, const double *x, double *y) {
#pragma omp parallel for schedule(static)
  for (i64 r = 0; r < rows; r++) {
    double v = 0.0;
    for (int s = 0; s < S; s++)
      if (col[r * S + s] >= 0) v += val[r * S + s] * x[col[r * S + s]];
    y[r] = v;
  }
}

void oracle_moments(i64 nodes, int A, const double *w, const double *psi, double *phi) {
#pragma omp parallel for schedule(static)
  for (i64 i = 0; i < nodes; i++) {
    double v = 0.0;
    for (int a = 0; a < A; a++) v += psi[i * A + a] * w[a];
    phi[i] = v;
  }
}

/* y(node, a) -= sigma_s(cell)/W * phi(node), phi = sum_b w_b x(node, b) */
void oracle_scatter_add(i64 nodes, int nb, int A, const double *w, double sum_w, const double *sigs,
                        const double *x, double *y) {
#pragma omp parallel for schedule(static)
  for (i64 i = 0; i < nodes; i++) {
    double phi = 0.0;
    for (int a = 0; a < A; a++) phi += x[i * A + a] * w[a];
    phi *= sigs[i / nb] / sum_w;
    for (int a = 0; a < A; a++) y[i * A + a] -= phi;
  }
}

/* b(node, a) += scale(cell) * phi(node) for every angle - the isotropic
 * broadcast of a group transfer (src/multigroupk.kokkos.cxx:332-359) with
 * scale = sigma_s(g' -> g)/W, or of a source with phi the basis-0 indicator */
void oracle_broadcast_add(i64 nodes, int nb, int A, const double *scale, const double *phi, double *b) {
#pragma omp parallel for schedule(static)
  for (i64 i = 0; i < nodes; i++) {
    const double v = scale[i / nb] * phi[i];
    for (int a = 0; a < A; a++) b[i * A + a] += v;
  }
}

/* ------------------------------------------------------------------------
 * Element blocks read off the assembled rows and inverted by Gauss-Jordan with
 * partial pivoting (src/block_inversek.kokkos.cxx:83-150); apply y = inv x
 * (:206-222). Returns the number of singular blocks; their inverse is zero.
 * ------------------------------------------------------------------------ */
i64 oracle_block_setup(int nb, i64 cells, int A, int S, const i64 *col, const double *val, double *inv) {
  i64 singular = 0;
#pragma omp parallel for schedule(static) reduction(+ : singular)
  for (i64 b = 0; b < cells * A; b++) {
    const i64 c = b / A;
    const int a = (int)(b % A);
    double m[4][4], w[4][4];
    for (int i = 0; i < nb; i++) {
      for (int j = 0; j < nb; j++) {
        m[i][j] = 0.0;
        w[i][j] = i == j ? 1.0 : 0.0;
      }
      const i64 r = (c * nb + i) * A + a;
      for (int s = 0; s < S; s++) {
        const i64 cl = col[r * S + s];
        if (cl < 0 || cl % A != a || cl / ((i64)nb * A) != c) continue;
        m[i][(cl / A) % nb] += val[r * S + s];
      }
    }
    int ok = 1;
    for (int k = 0; k < nb && ok; k++) {
      int piv = k;
      for (int i = k + 1; i < nb; i++)
        if (fabs(m[i][k]) > fabs(m[piv][k])) piv = i;
      if (fabs(m[piv][k]) == 0.0) {
        ok = 0;
        break;
      }
      if (piv != k)
        for (int j = 0; j < nb; j++) {
          double t = m[k][j];
          m[k][j] = m[piv][j];
          m[piv][j] = t;
          t = w[k][j];
          w[k][j] = w[piv][j];
          w[piv][j] = t;
        }
      const double ip = 1.0 / m[k][k];
      for (int j = 0; j < nb; j++) {
        m[k][j] *= ip;
        w[k][j] *= ip;
      }
      for (int i = 0; i < nb; i++) {
        if (i == k) continue;
        const double f = m[i][k];
        if (f == 0.0) continue;
        for (int j = 0; j < nb; j++) {
          m[i][j] -= f * m[k][j];
          w[i][j] -= f * w[k][j];
        }
      }
    }
    if (!ok) singular++;
    for (int i = 0; i < nb; i++)
      for (int j = 0; j < nb; j++) inv[(b * nb + i) * nb + j] = ok ? w[i][j] : 0.0;
  }
  return singular;
}

void oracle_block_apply(int nb, i64 cells, int A, const double *inv, const double *x, double *y) {
#pragma omp parallel for schedule(static)
  for (i64 b = 0; b < cells * A; b++) {
    const i64 c = b / A;
    const int a = (int)(b % A);
    double v[4];
    for (int j = 0; j < nb; j++) v[j] = x[(c * nb + j) * A + a];
    for (int i = 0; i < nb; i++) {
      double s = 0.0;
      for (int j = 0; j < nb; j++) s += inv[(b * nb + i) * nb + j] * v[j];
      y[(c * nb + i) * A + a] = s;
    }
  }
}
