/*
 * nrg_oracle.c -- FP64 single-thread CPU restatement of the TwoChNRG
 * iterative-diagonalisation path (OneChQS Anderson through NRG_main).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity checker for the HIP
 * library and the timed "port" CPU baseline.  The product library never links,
 * loads or calls it.  It follows the reference's loops literally (linear
 * block searches, dense row-major GEMMs, unblocked Householder + implicit QL)
 * so that the data layouts and orderings it produces are the reference's.
 *
 * Pinning (see tests/test_oracle.py): golden run Runs/run1 (636 levels,
 * block labels, Emin/Ecut, <n_d>,<n_d^2>,<S_d^2>), closed forms (H_{-1},
 * chi_N), brute-force Fock-space exact diagonalisation, U=0 single-particle
 * sums, LAPACK (scipy) eigenvalues, CG tables.
 */
#include "nrg_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ------------------------------------------------------------------ */
/* growable arrays                                                      */
/* ------------------------------------------------------------------ */
typedef struct { double *a; int n, cap; } dvec;
typedef struct { int *a; int n, cap; } ivec;

static void dres(dvec *v, int cap) {
  if (cap > v->cap) {
    int c = v->cap ? v->cap : 16;
    while (c < cap) c *= 2;
    v->a = (double *)realloc(v->a, sizeof(double) * (size_t)c);
    v->cap = c;
  }
}
static void ires(ivec *v, int cap) {
  if (cap > v->cap) {
    int c = v->cap ? v->cap : 16;
    while (c < cap) c *= 2;
    v->a = (int *)realloc(v->a, sizeof(int) * (size_t)c);
    v->cap = c;
  }
}
static void dpush(dvec *v, double x) { dres(v, v->n + 1); v->a[v->n++] = x; }
static void ipush(ivec *v, int x) { ires(v, v->n + 1); v->a[v->n++] = x; }
static void dcopy_v(dvec *dst, const dvec *src) {
  dres(dst, src->n); if (src->n) memcpy(dst->a, src->a, sizeof(double) * (size_t)src->n); dst->n = src->n;
}
static void icopy_v(ivec *dst, const ivec *src) {
  ires(dst, src->n); if (src->n) memcpy(dst->a, src->a, sizeof(int) * (size_t)src->n); dst->n = src->n;
}
static void derase(dvec *v, int b, int e) { /* erase [b,e) */
  if (e <= b) return;
  memmove(v->a + b, v->a + e, sizeof(double) * (size_t)(v->n - e));
  v->n -= (e - b);
}
static void ierase(ivec *v, int b, int e) {
  if (e <= b) return;
  memmove(v->a + b, v->a + e, sizeof(int) * (size_t)(v->n - e));
  v->n -= (e - b);
}

/* ------------------------------------------------------------------ */
/* CNRGarray / CNRGbasisarray / CNRGmatrix subset                       */
/* include/NRGclasses.hpp:47-759                                        */
/* ------------------------------------------------------------------ */
typedef struct {
  int nshell, nq;
  dvec qn;      /* QNumbers */
  ivec bbe;     /* BlockBegEnd */
  dvec en;      /* dEn */
  dvec ev;      /* dEigVec */
  ivec kept;    /* Kept */
  ivec itype;   /* CNRGbasisarray::iType */
  ivec stcf;    /* CNRGbasisarray::StCameFrom */
  ivec bbe_bc;  /* BlockBegEndBC */
  ivec bbe_ev;  /* BlockBegEndEigVec */
  int sz;       /* (Q, Sz) blocks without SU(2): degeneracy 1, no Clebsch-Gordan factors */
} arr_t;

typedef struct {
  arr_t b;          /* synced array part */
  dvec mat;         /* MatEl */
  ivec map;         /* MatBlockMap */
  ivec mbe;         /* MatBlockBegEnd */
  int upper, need_old, calc_avg;
  int (*check)(const arr_t *, int, int);
  double (*calc)(const arr_t *, const arr_t *, int, int);
} mat_t;

static void arr_free(arr_t *A) {
  free(A->qn.a); free(A->bbe.a); free(A->en.a); free(A->ev.a); free(A->kept.a);
  free(A->itype.a); free(A->stcf.a); free(A->bbe_bc.a); free(A->bbe_ev.a);
  memset(A, 0, sizeof(*A));
}
static void mat_free(mat_t *M) {
  arr_free(&M->b); free(M->mat.a); free(M->map.a); free(M->mbe.a);
  memset(M, 0, sizeof(*M));
}
/* CNRGarray::ClearAll  src/NRGarrayClass.cpp (ClearAll) + CNRGbasisarray::ClearAll :1150-1161 */
static void arr_clear(arr_t *A) {
  A->qn.n = A->bbe.n = A->en.n = A->ev.n = A->kept.n = 0;
  A->itype.n = A->stcf.n = A->bbe_bc.n = A->bbe_ev.n = 0;
}
static int nblocks(const arr_t *A) { return A->nq ? A->qn.n / A->nq : 0; }
static int blim(const arr_t *A, int b, int w) { return b < 0 ? 0 : A->bbe.a[2 * b + w]; }
static int bsize(const arr_t *A, int b) { return A->bbe.a[2 * b + 1] - A->bbe.a[2 * b] + 1; }
static int nstates(const arr_t *A) { return A->bbe.a[A->bbe.n - 1] - A->bbe.a[0] + 1; }
static double getqn(const arr_t *A, int b, int q) { return A->qn.n ? A->qn.a[b * A->nq + q] : 0.0; }
/* GetBlockFromSt  src/NRGarrayClass.cpp:374-390 (linear) */
static int blk_from_st(const arr_t *A, int ist, int *pos) {
  int b = 0;
  while (ist > A->bbe.a[2 * b + 1]) b++;
  *pos = ist - A->bbe.a[2 * b];
  return b;
}
static double qn_from_st(const arr_t *A, int ist, int q) {
  int pos; int b = blk_from_st(A, ist, &pos); return getqn(A, b, q);
}
/* dEqual  src/dEqual.cpp:6-10 */
static int dEqual(double a, double b) { return fabs(a - b) < 1e-20; }
/* GetBlockFromQNumbers  src/NRGarrayClass.cpp:514-542: first block whose QNs all dEqual */
static int blk_from_qn(const arr_t *A, const double *q) {
  int nb = nblocks(A);
  for (int b = 0; b < nb; b++) {
    int ok = 1;
    for (int i = 0; i < A->nq; i++) if (!dEqual(A->qn.a[b * A->nq + i], q[i])) { ok = 0; break; }
    if (ok) return b;
  }
  return -1;
}
/* GetBlockLimitEigv: CNRGarray sums n^2 (:400-415); CNRGbasisarray uses BlockBegEndEigVec (:1550-1564) */
static int blim_ev(const arr_t *A, int b) {
  if (A->bbe_ev.n) return A->bbe_ev.a[2 * b];
  int pos = 0;
  for (int i = 0; i < b; i++) { int n = bsize(A, i); pos += n * n; }
  return pos;
}
static int bsize_bc(const arr_t *A, int b) {
  if (A->bbe_bc.n == 0) return bsize(A, b);
  return A->bbe_bc.a[2 * b + 1] - A->bbe_bc.a[2 * b] + 1;
}

/* CNRGmatrix::FindMatBlock  src/NRGmatrixClass.cpp:33-70: first (b1,b2) pair at an even slot */
static int find_mat_block(const mat_t *M, int b1, int b2) {
  for (int i = 0; i + 1 < M->map.n; i += 2)
    if (M->map.a[i] == b1 && M->map.a[i + 1] == b2) return i / 2;
  return -1;
}
static int ij2rUpTr(int n, int i, int j) { /* src/ij2r.cpp:63-68 */
  if (i > j || i < 0 || j < 0) return 0;
  return i * n - i * (i - 1) / 2 + (j - i);
}
/* GetBlockMatEl  src/NRGmatrixClass.cpp:120-150 */
static double get_block_matel(const mat_t *M, int b1, int b2, int i, int j) {
  int im = find_mat_block(M, b1, b2);
  if (im < 0 || i < 0 || j < 0) return 0.0;
  int i0 = M->mbe.a[2 * im];
  int n1 = bsize(&M->b, b1), n2 = bsize(&M->b, b2);
  int r = n2 * i + j; (void)n1;
  if (M->upper && b1 == b2) r = (i > j) ? ij2rUpTr(n1, j, i) : ij2rUpTr(n1, i, j);
  return M->mat.a[i0 + r];
}
/* GetMatEl  :160-170 */
static double get_matel(const mat_t *M, int ist, int jst) {
  int pi, pj;
  int b1 = blk_from_st(&M->b, ist, &pi), b2 = blk_from_st(&M->b, jst, &pj);
  return get_block_matel(M, b1, b2, pi, pj);
}
/* SyncNRGarray(CNRGbasisarray)  include/NRGclasses.hpp:446-456 */
static void mat_sync(mat_t *M, const arr_t *A) {
  M->b.nshell = A->nshell; M->b.nq = A->nq;
  dcopy_v(&M->b.qn, &A->qn); icopy_v(&M->b.bbe, &A->bbe); icopy_v(&M->b.kept, &A->kept);
  icopy_v(&M->b.itype, &A->itype);
}
static void mat_clear(mat_t *M) {  /* CNRGmatrix::ClearAll :416-425 */
  M->b.qn.n = M->b.bbe.n = M->b.en.n = M->b.ev.n = M->b.kept.n = 0;
  M->mat.n = M->map.n = M->mbe.n = 0;
}
static void mat_copy(mat_t *dst, const mat_t *src) {
  dst->b.nshell = src->b.nshell; dst->b.nq = src->b.nq;
  dcopy_v(&dst->b.qn, &src->b.qn); icopy_v(&dst->b.bbe, &src->b.bbe);
  icopy_v(&dst->b.kept, &src->b.kept); icopy_v(&dst->b.itype, &src->b.itype);
  dcopy_v(&dst->mat, &src->mat); icopy_v(&dst->map, &src->map); icopy_v(&dst->mbe, &src->mbe);
  dst->upper = src->upper; dst->need_old = src->need_old; dst->calc_avg = src->calc_avg;
  dst->check = src->check; dst->calc = src->calc;
}
/* SetDiagonalMatrix  src/NRGmatrixClass.cpp:1727-1769 */
static void mat_set_diag(mat_t *M, double el) {
  M->map.n = M->mbe.n = M->mat.n = 0;
  int i1 = 0;
  for (int b = 0; b < nblocks(&M->b); b++) {
    ipush(&M->map, b); ipush(&M->map, b); ipush(&M->mbe, i1);
    int n = bsize(&M->b, b);
    for (int i = 0; i < n; i++) for (int j = 0; j < n; j++) { dpush(&M->mat, i == j ? el : 0.0); i1++; }
    ipush(&M->mbe, i1 - 1);
  }
}
/* PushBlockMatEl  src/NRGmatrixClass.cpp:267-290 (non-upper) */
static void mat_put(mat_t *M, int b1, int b2, int i, int j, double el) {
  int im = find_mat_block(M, b1, b2);
  if (im < 0) return;
  M->mat.a[M->mbe.a[2 * im] + bsize(&M->b, b2) * i + j] = el;
}

/* ------------------------------------------------------------------ */
/* Clebsch-Gordan  src/CGordan.cpp:13-120 (Racah formula, exact factorials) */
/* ------------------------------------------------------------------ */
static double fact(int n) {
  static double tab[171];
  static int init = 0;
  if (!init) { tab[0] = 1.0; for (int i = 1; i < 171; i++) tab[i] = tab[i - 1] * (double)i; init = 1; }
  if (n < 0 || n > 170) return 0.0;
  return tab[n];
}
static int triang(double J1, double J2, double J) {  /* TriangIneq :13-17 */
  return !((J - (J1 + J2) > 1.0e-10) || (J - fabs(J1 - J2) < -1.0e-10));
}
double orc_cgordan(double J1, double M1, double J2, double M2, double J, double M) {
  if (J1 < 0.0 || J2 < 0.0 || J < 0.0) return 0.0;
  if (fabs(M1) > J1 || fabs(M2) > J2 || fabs(M) > J) return 0.0;
  if (fabs(M - (M1 + M2)) > 1.0e-10) return 0.0;
  if (J > J1 + J2 || J < fabs(J1 - J2)) return 0.0;
  double a5 = fact((int)(J1 + J2 - J)) * fact((int)(J1 - J2 + J)) * fact((int)(-J1 + J2 + J)) *
              (2 * J + 1) / fact((int)(J1 + J2 + J + 1));
  double rho = a5 > 0.0 ? sqrt(a5) : 0.0;
  double a6 = fact((int)(J + M)) * fact((int)(J - M)) * fact((int)(J1 + M1)) * fact((int)(J1 - M1)) *
              fact((int)(J2 + M2)) * fact((int)(J2 - M2));
  double sigma = a6 > 0.0 ? sqrt(a6) : 0.0;
  int r1 = (int)(J1 - M1), r2 = (int)(J2 + M2), r3 = (int)(J1 + J2 - J);
  int rmax = r1 < r2 ? r1 : r2; if (r3 < rmax) rmax = r3;
  int m1 = -(int)(J - J2 + M1), m2 = -(int)(J - J1 - M2);
  int rmin = m1 > m2 ? m1 : m2; if (rmin < 0) rmin = 0;
  double tau = 0.0;
  for (int r = rmin; r <= rmax; r++) {
    double d = fact((int)(J1 - M1) - r) * fact((int)(J2 + M2) - r) * fact((int)(J - J2 + M1) + r) *
               fact((int)(J - J1 - M2) + r) * fact((int)(J1 + J2 - J) - r) * fact(r);
    tau += ((r & 1) ? -1.0 : 1.0) * (1.0 / d);
  }
  return rho * sigma * tau;
}

/* ------------------------------------------------------------------ */
/* Symmetric eigensolver: the dspev_/dsyev_ calls of DiagBlock           */
/* (src/NRGmatrixClass.cpp:566-696).  Householder reduction to           */
/* tridiagonal form with explicit accumulation of Q, then implicit QL    */
/* with Wilkinson shifts on (d,e) rotating the columns of Q; eigenpairs  */
/* sorted ascending and returned as ROWS (row i = eigenvector i), which  */
/* is the layout DiagBlock copies into dEigVec.                          */
/* ------------------------------------------------------------------ */
static void eigh_dense(int n, double *a /* n*n row-major, overwritten by Q */, double *w, double *z) {
  double *d = (double *)calloc((size_t)n, sizeof(double));
  double *e = (double *)calloc((size_t)n, sizeof(double));
  double *v = (double *)calloc((size_t)n, sizeof(double));
  double *p = (double *)calloc((size_t)n, sizeof(double));
  /* Q accumulated as product H_0 H_1 ... ; store reflectors then build Q */
  double *vs = (double *)calloc((size_t)n * (size_t)n, sizeof(double));
  double *tau = (double *)calloc((size_t)n, sizeof(double));
#define A(i, j) a[(size_t)(i) * n + (j)]
  for (int k = 0; k + 2 < n; k++) {
    /* reflector annihilating A(k+2:n, k) */
    double nrm = 0.0;
    for (int i = k + 1; i < n; i++) nrm += A(i, k) * A(i, k);
    nrm = sqrt(nrm);
    double x0 = A(k + 1, k);
    if (nrm == 0.0) { tau[k] = 0.0; continue; }
    double alpha = (x0 > 0) ? -nrm : nrm;
    /* v = x - alpha e1, H = I - 2 v v^T / (v^T v) */
    for (int i = 0; i < n; i++) v[i] = 0.0;
    v[k + 1] = x0 - alpha;
    for (int i = k + 2; i < n; i++) v[i] = A(i, k);
    double vtv = 0.0;
    for (int i = k + 1; i < n; i++) vtv += v[i] * v[i];
    if (vtv == 0.0) { tau[k] = 0.0; continue; }
    double t = 2.0 / vtv;
    tau[k] = t;
    for (int i = k + 1; i < n; i++) vs[(size_t)k * n + i] = v[i];
    /* A <- H A H on the trailing block (rows/cols k..n-1) */
    /* p = t * A v */
    for (int i = k; i < n; i++) {
      double s = 0.0;
      for (int j = k + 1; j < n; j++) s += A(i, j) * v[j];
      p[i] = t * s;
    }
    double vp = 0.0;
    for (int i = k + 1; i < n; i++) vp += v[i] * p[i];
    double K = 0.5 * t * vp;
    /* q = p - K v ; A <- A - v q^T - q v^T */
    for (int i = k; i < n; i++) p[i] -= K * v[i];
    for (int i = k; i < n; i++)
      for (int j = k; j < n; j++) A(i, j) -= v[i] * p[j] + p[i] * v[j];
    /* exact zeros below the subdiagonal */
    A(k + 1, k) = alpha; A(k, k + 1) = alpha;
    for (int i = k + 2; i < n; i++) { A(i, k) = 0.0; A(k, i) = 0.0; }
  }
  for (int i = 0; i < n; i++) d[i] = A(i, i);
  for (int i = 0; i + 1 < n; i++) e[i] = A(i + 1, i);
  e[n - 1] = 0.0;
  /* Q = H_0 H_1 ... H_{n-3}: apply in reverse to the identity */
  for (int i = 0; i < n; i++) for (int j = 0; j < n; j++) A(i, j) = (i == j) ? 1.0 : 0.0;
  for (int k = n - 3; k >= 0; k--) {
    if (tau[k] == 0.0) continue;
    const double *vk = vs + (size_t)k * n;
    for (int j = k + 1; j < n; j++) {
      double s = 0.0;
      for (int i = k + 1; i < n; i++) s += vk[i] * A(i, j);
      s *= tau[k];
      for (int i = k + 1; i < n; i++) A(i, j) -= s * vk[i];
    }
  }
  /* implicit QL on (d,e); rotations applied to columns of Q (=A) */
  for (int l = 0; l < n; l++) {
    int iter = 0;
    for (;;) {
      int m;
      for (m = l; m + 1 < n; m++) {
        double dd = fabs(d[m]) + fabs(d[m + 1]);
        if (fabs(e[m]) <= 1e-300 || fabs(e[m]) + dd == dd) break;
      }
      if (m == l) break;
      if (++iter > 60) break;
      double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
      double r = hypot(g, 1.0);
      g = d[m] - d[l] + e[l] / (g + (g >= 0 ? fabs(r) : -fabs(r)));
      double s = 1.0, c = 1.0, pp = 0.0;
      int i;
      for (i = m - 1; i >= l; i--) {
        double f = s * e[i], b = c * e[i];
        r = hypot(f, g);
        e[i + 1] = r;
        if (r == 0.0) { d[i + 1] -= pp; e[m] = 0.0; break; }
        s = f / r; c = g / r;
        g = d[i + 1] - pp;
        r = (d[i] - g) * s + 2.0 * c * b;
        pp = s * r;
        d[i + 1] = g + pp;
        g = c * r - b;
        for (int k = 0; k < n; k++) {
          double t1 = A(k, i + 1);
          A(k, i + 1) = s * A(k, i) + c * t1;
          A(k, i) = c * A(k, i) - s * t1;
        }
      }
      if (r == 0.0 && i >= l) continue;
      d[l] -= pp; e[l] = g; e[m] = 0.0;
    }
  }
  /* sort ascending (stable selection) and emit rows */
  int *idx = (int *)malloc(sizeof(int) * (size_t)n);
  for (int i = 0; i < n; i++) idx[i] = i;
  for (int i = 0; i < n; i++) {
    int kmin = i;
    for (int j = i + 1; j < n; j++) if (d[idx[j]] < d[idx[kmin]]) kmin = j;
    int t = idx[i]; idx[i] = idx[kmin]; idx[kmin] = t;
  }
  for (int i = 0; i < n; i++) {
    w[i] = d[idx[i]];
    for (int k = 0; k < n; k++) z[(size_t)i * n + k] = A(k, idx[i]);
  }
#undef A
  free(idx); free(d); free(e); free(v); free(p); free(vs); free(tau);
}

void orc_eigh_full(int n, const double *a, double *w, double *z) {
  if (n <= 0) return;
  double *t = (double *)malloc(sizeof(double) * (size_t)n * (size_t)n);
  memcpy(t, a, sizeof(double) * (size_t)n * (size_t)n);
  eigh_dense(n, t, w, z);
  free(t);
}
void orc_eigh_packed(int n, const double *ap, double *w, double *z) {
  if (n <= 0) return;
  double *t = (double *)malloc(sizeof(double) * (size_t)n * (size_t)n);
  for (int i = 0; i < n; i++)
    for (int j = i; j < n; j++) { double x = ap[ij2rUpTr(n, i, j)]; t[(size_t)i * n + j] = x; t[(size_t)j * n + i] = x; }
  eigh_dense(n, t, w, z);
  free(t);
}

/* ------------------------------------------------------------------ */
/* SetE0zero / SetDegen / Ecut / SetKept  src/NRGarrayClass.cpp:234-491, 661-691 */
/* ------------------------------------------------------------------ */
static double set_e0zero(arr_t *A) {
  double emin = A->en.a[0];
  for (int i = 1; i < A->en.n; i++) if (A->en.a[i] < emin) emin = A->en.a[i];
  for (int i = 0; i < A->en.n; i++) A->en.a[i] -= emin;
  return emin;
}
void orc_setdegen(int n, double *en, double myeps) {
  for (int ii = 0; ii < n; ii++)
    for (int jj = ii + 1; jj < n; jj++)
      if (fabs(en[jj] - en[ii]) < myeps) {
        if (en[jj] > en[ii]) en[jj] = en[ii]; else en[ii] = en[jj];
      }
}
static int cmpd(const void *x, const void *y) {
  double a = *(const double *)x, b = *(const double *)y;
  return (a > b) - (a < b);
}
/* The 1% window loop reads Eaux[Ncut+ii] with the bound ii<size only
 * (:682-684), i.e. possibly past the end.  Here an index past the end stops
 * the window as if a level far above Ecut were there, which keeps every state. */
double orc_ecut(int n, const double *en, int ncut) {
  double *e = (double *)malloc(sizeof(double) * (size_t)n);
  memcpy(e, en, sizeof(double) * (size_t)n);
  qsort(e, (size_t)n, sizeof(double), cmpd);
  double aux;
  if (ncut >= n - 1) aux = e[n - 1];
  else {
    aux = e[ncut - 1];
    double aux2 = e[ncut];
    int ii = 1;
    while (aux2 > 0.0 && fabs((aux2 - aux) / aux) < 0.01 && ii < n) {
      if (ncut + ii >= n) { aux2 = HUGE_VAL; ii++; break; }
      aux2 = e[ncut + ii]; ii++;
    }
    if (ii > 1) aux = (ncut + ii - 2 < n) ? e[ncut + ii - 2] : e[n - 1];
  }
  free(e);
  return aux;
}
static double set_kept(arr_t *A, int ncut, double myeps, int *nkept) {
  orc_setdegen(A->en.n, A->en.a, myeps);
  double ec = orc_ecut(A->en.n, A->en.a, ncut);
  A->kept.n = 0;
  int nk = 0;
  for (int i = 0; i < A->en.n; i++) { int k = A->en.a[i] <= ec; ipush(&A->kept, k); nk += k; }
  if (nkept) *nkept = nk;
  return ec;
}

/* ------------------------------------------------------------------ */
/* CutStates  src/CutStates.cpp:5-96 with RemoveBlock / RemoveStatesFromBlock */
/* (src/NRGarrayClass.cpp:1307-1548)                                    */
/* ------------------------------------------------------------------ */
static void set_bbe_ev(arr_t *A) {  /* SetBlockBegEndEigVec :1285-1305 */
  A->bbe_ev.n = 0;
  int nb = A->bbe_bc.n / 2, aux = 0;
  for (int b = 0; b < nb; b++) {
    ipush(&A->bbe_ev, aux);
    int n = bsize_bc(A, b);
    aux += n * n - 1;
    ipush(&A->bbe_ev, aux);
    aux++;
  }
}
static void remove_block(arr_t *A, int ib) {
  int b0 = blim(A, ib, 0), b1 = blim(A, ib, 1);
  derase(&A->en, b0, b1 + 1);
  ierase(&A->stcf, b0, b1 + 1);
  ierase(&A->kept, b0, b1 + 1);
  derase(&A->ev, A->bbe_ev.a[2 * ib], A->bbe_ev.a[2 * ib + 1] + 1);
  derase(&A->qn, A->nq * ib, A->nq * (ib + 1));
  int nb = nblocks(A);
  for (int b = ib; b < nb; b++) {
    int sn = A->bbe.a[2 * (b + 1) + 1] - A->bbe.a[2 * (b + 1)] + 1;
    A->bbe.a[2 * b + 1] = A->bbe.a[2 * b] + sn - 1;
    A->bbe.a[2 * b + 2] = A->bbe.a[2 * b] + sn;
    int bn = A->bbe_bc.a[2 * (b + 1) + 1] - A->bbe_bc.a[2 * (b + 1)] + 1; /* BC size of next */
    (void)bn;
    int sn2 = A->bbe_ev.a[2 * (b + 1) + 1] - A->bbe_ev.a[2 * (b + 1)] + 1;
    /* The reference recomputes the eigvec range from Size_next^2 (:1340-1344):
       at cut time every block is still square, so the two agree. */
    A->bbe_ev.a[2 * b + 1] = A->bbe_ev.a[2 * b] + sn2 - 1;
    A->bbe_ev.a[2 * b + 2] = A->bbe_ev.a[2 * b] + sn2;
  }
  A->bbe.n -= 2; A->bbe_ev.n -= 2;
  ierase(&A->bbe_bc, 2 * ib, 2 * (ib + 1));
}
static void remove_states_from_block(arr_t *A, int ib, int ist1, int ist2) {
  int ngone = ist2 - ist1 + 1;
  int p0, p1;
  if (blk_from_st(A, ist1, &p0) != ib || blk_from_st(A, ist2, &p1) != ib) return;
  int nbc = bsize_bc(A, ib);
  if (A->bbe.a[2 * ib] == A->bbe.a[2 * ib + 1]) { remove_block(A, ib); return; }
  derase(&A->en, ist1, ist2 + 1);
  ierase(&A->stcf, ist1, ist2 + 1);
  ierase(&A->kept, ist1, ist2 + 1);
  int off = A->bbe_ev.a[2 * ib] + nbc * p0;
  derase(&A->ev, off, off + nbc * ngone);
  A->bbe.a[2 * ib + 1] -= ngone;
  A->bbe_ev.a[2 * ib + 1] -= nbc * ngone;
  for (int b = ib + 1; b < nblocks(A); b++) {
    A->bbe.a[2 * b] -= ngone; A->bbe.a[2 * b + 1] -= ngone;
    A->bbe_ev.a[2 * b] -= nbc * ngone; A->bbe_ev.a[2 * b + 1] -= nbc * ngone;
  }
}
static double cut_states(const arr_t *Ain, arr_t *Ac, int ncut, double myeps) {
  arr_clear(Ac);
  Ac->nshell = Ain->nshell; Ac->nq = Ain->nq;
  dcopy_v(&Ac->qn, &Ain->qn); dcopy_v(&Ac->en, &Ain->en); dcopy_v(&Ac->ev, &Ain->ev);
  icopy_v(&Ac->kept, &Ain->kept); icopy_v(&Ac->bbe, &Ain->bbe);
  Ac->stcf.n = 0;
  for (int i = 0; i < Ac->en.n; i++) ipush(&Ac->stcf, i);
  icopy_v(&Ac->bbe_bc, &Ain->bbe);
  set_bbe_ev(Ac);
  if (ncut >= nstates(Ain)) return HUGE_VAL;
  double ec = orc_ecut(Ain->en.n, Ain->en.a, ncut);
  set_kept(Ac, ncut, myeps, NULL);
  for (int b = 0; b < nblocks(Ac); b++) {
    int b0 = blim(Ac, b, 0), b1 = blim(Ac, b, 1);
    double emin = Ac->en.a[b0];
    for (int i = b0; i <= b1; i++) if (Ac->en.a[i] < emin) emin = Ac->en.a[i];
    if (emin > ec) { remove_block(Ac, b); b--; }
  }
  for (int b = 0; b < nblocks(Ac); b++) {
    int b0 = blim(Ac, b, 0), b1 = blim(Ac, b, 1);
    double emax = Ac->en.a[b0];
    for (int i = b0; i <= b1; i++) if (Ac->en.a[i] > emax) emax = Ac->en.a[i];
    if (emax > ec) {
      int ip = b0;  /* GetBlockEcutPos :450-459 */
      while (ip <= b1 && Ac->en.a[ip] <= ec) ip++;
      remove_states_from_block(Ac, b, ip, b1);
    }
  }
  return ec;
}

/* ------------------------------------------------------------------ */
/* OneChQS model plugin                                                 */
/* ------------------------------------------------------------------ */
/* OneChQS_SetSingleSite  src/OneChQS_SetSingleSite.cpp */
static void onechqs_site(arr_t *S) {
  arr_clear(S);
  S->nq = 3; S->nshell = 0;
  const double q[12] = {-1, 0, 0, 0, 0.5, 0.5, 0, 0.5, -0.5, 1, 0, 0};
  for (int i = 0; i < 12; i++) dpush(&S->qn, q[i]);
  for (int i = 0; i < 4; i++) { ipush(&S->bbe, i); ipush(&S->bbe, i); ipush(&S->itype, i); }
}
/* OneCh_fd_table  src/OpTables.cpp:8-33 */
static double onech_fd(int sigma, int ti, int tj) {
  if (sigma == 1) {
    if (ti == 1 && tj == 0) return 1.0;
    if (ti == 3 && tj == 2) return 1.0;
    return 0.0;
  }
  if (sigma == -1) {
    if (ti == 2 && tj == 0) return 1.0;
    if (ti == 3 && tj == 1) return -1.0;
    return 0.0;
  }
  return 0.0;
}
/* Diag_check  src/OpMatRules.cpp:12-24 */
static int diag_check(const arr_t *A, int b1, int b2) { (void)A; return b1 == b2; }
/* OneChQS_cd_check  src/OpMatRules.cpp:106-125 */
static int onechqs_cd_check(const arr_t *A, int b1, int b2) {
  double Qi = getqn(A, b1, 0), Si = getqn(A, b1, 1), Qj = getqn(A, b2, 0), Sj = getqn(A, b2, 1);
  return dEqual(Qj, Qi + 1.0) && (dEqual(Sj, Si + 0.5) || dEqual(Sj, Si - 0.5));
}
/* ImpOnly_MatEl  src/OpMatRules.cpp:265-280 */
static double imponly_matel(const arr_t *B, const arr_t *S, int i, int j) {
  (void)S; return B->itype.a[i] == B->itype.a[j] ? 1.0 : 0.0;
}
/* OneChQS_fN_MatEl  src/OpMatRules.cpp:1167-1247 */
static double onechqs_fn_matel(const arr_t *B, const arr_t *S, int ist, int jst) {
  double Si = qn_from_st(B, ist, 1), Sj = qn_from_st(B, jst, 1);
  int ti = B->itype.a[ist], tj = B->itype.a[jst];
  if (B->stcf.a[ist] != B->stcf.a[jst]) return 0.0;
  double Qti = getqn(S, ti, 0), Sti = getqn(S, ti, 1), Szti = getqn(S, ti, 2);
  double Qtj = getqn(S, tj, 0), Stj = getqn(S, tj, 1), Sztj = getqn(S, tj, 2);
  double qi[3] = {Qti, Sti, Szti}, qj[3] = {Qtj, Stj, Sztj};
  double Scfi = Si - Szti, Scfj = Sj - Sztj;
  double acc = 0.0, cg1 = 0.0;
  double sigma = Sj - Si;
  for (double sz = -Stj; sz <= Stj; sz += 1.0) {
    qj[2] = sz;
    int ssj = blk_from_qn(S, qj);
    qi[2] = sz - sigma;
    int ssi = blk_from_qn(S, qi);
    double cg0 = orc_cgordan(Scfi, Si - sz + sigma, Sti, sz - sigma, Si, Si);
    cg1 = orc_cgordan(Scfj, Sj - sz, Stj, sz, Sj, Sj);
    acc += onech_fd((int)(sigma / 0.5), ssj, ssi) * cg0 * cg1;
  }
  double cgr = orc_cgordan(Si, Si, 0.5, sigma, Sj, Sj);
  if (!dEqual(cgr, 0.0)) return acc / cgr;
  return 0.0;  /* auxBasis[1] is never accumulated (i1 resets to 0 each pass) */
}
/* OneChQS_cd_MatEl  src/OpMatRules.cpp:1249-1332 */
static double onechqs_cd_matel(const arr_t *B, const arr_t *S, int ist, int jst) {
  double Si = qn_from_st(B, ist, 1), Sj = qn_from_st(B, jst, 1);
  int ti = B->itype.a[ist], tj = B->itype.a[jst];
  double Qti = getqn(S, ti, 0), Sti = getqn(S, ti, 1), Szti = getqn(S, ti, 2);
  double Qtj = getqn(S, tj, 0), Stj = getqn(S, tj, 1), Sztj = getqn(S, tj, 2);
  double qi[3] = {Qti, Sti, Szti}, qj[3] = {Qtj, Stj, Sztj};
  double Scfi = Si - Szti, Scfj = Sj - Sztj;
  double sigma = Sj - Si;
  double cg3 = orc_cgordan(Si, Si, 0.5, sigma, Sj, Sj);
  if (dEqual(cg3, 0.0)) {
    sigma = Si - Sj;
    cg3 = orc_cgordan(Si, Si, 0.5, sigma, Sj, Sj);
    if (dEqual(cg3, 0.0)) return 0.0;
  }
  double fb = 0.0;
  for (double szcfj = -Scfj; szcfj <= Scfj; szcfj += 1.0) {
    double szcfi = szcfj - sigma;
    double sztj = Sj - szcfj, szti = Si - szcfi;
    qj[2] = sztj; int ssj = blk_from_qn(S, qj);
    qi[2] = szti; int ssi = blk_from_qn(S, qi);
    double c0 = orc_cgordan(Scfi, szcfi, Sti, szti, Si, Si);
    double c1 = orc_cgordan(Scfj, szcfj, Stj, sztj, Sj, Sj);
    double c2 = orc_cgordan(Scfi, szcfi, 0.5, sigma, Scfj, szcfj);
    double fs = dEqual(Qtj, 0.0) ? -1.0 : 1.0;
    if (ssi == ssj) fb += c0 * c1 * c2 * fs;
  }
  if (fabs(cg3) > 1e-20) return fb / cg3;
  return 0.0;
}
/* Coupling part of OneChQS_HN_MatEl (src/OpMatRules.cpp:1460-1530) for
 * ist != jst of one (Q,S) block: returns CoefMatEl*FermiSign and the old
 * state pair (*pa, *pb) whose f element it multiplies.  structural = 0 is the
 * reference's rule (the (jst, ist) element with swapped site types when the
 * (ist, jst) one is zero, :1477-1488); structural = 1 takes the orientation
 * in which f stores the block pair, the same choice except where a stored
 * element happens to vanish (what a derivative of the element needs). */
static double hn_coupling(const arr_t *B, const arr_t *S, const mat_t *f, int ist, int jst, int structural,
                          int *pa, int *pb) {
  double Si = qn_from_st(B, ist, 1);
  int ti = B->itype.a[ist], tj = B->itype.a[jst];
  int a = B->stcf.a[ist], b = B->stcf.a[jst];
  int swap;
  if (structural) {
    int p1, p2;
    int b1 = blk_from_st(&f->b, a, &p1), b2 = blk_from_st(&f->b, b, &p2);
    swap = find_mat_block(f, b1, b2) < 0;
  } else {
    swap = dEqual(get_matel(f, a, b), 0.0);
  }
  int typep = ti, type = tj;
  if (swap) { typep = tj; type = ti; int x = a; a = b; b = x; }
  *pa = a; *pb = b;
  double Qtp = getqn(S, typep, 0), Stp = getqn(S, typep, 1), Sztp = getqn(S, typep, 2);
  double Qt = getqn(S, type, 0), St = getqn(S, type, 1), Szt = getqn(S, type, 2);
  double Scfp = Si - Sztp, Scf = Si - Szt;
  double fsign = dEqual(Qt, 0.0) ? -1.0 : 1.0;
  double coef = 0.0;
  double qp[3] = {Qtp, Stp, Sztp}, qq[3] = {Qt, St, Szt};
  for (double sz = -St; sz <= St; sz += 1.0) {
    qq[2] = sz;
    int ss = blk_from_qn(S, qq);
    for (int sg = -1; sg <= 1; sg += 2) {
      double ds = 0.5 * sg;
      double szcf = Si - sz, szcfp = szcf - ds, sztp = sz + ds;
      qp[2] = sz + ds;
      int ssp = blk_from_qn(S, qp);
      double c0 = orc_cgordan(Scf, szcf, St, sz, Si, Si);
      double c1 = orc_cgordan(Scfp, szcfp, Stp, sztp, Si, Si);
      double c2 = orc_cgordan(Scfp, szcfp, 0.5, ds, Scf, szcf);
      double c3 = onech_fd(sg, ssp, ss);
      coef += c0 * c1 * c2 * c3;
    }
  }
  return coef * fsign;
}
/* OneChQS_HN_MatEl  src/OpMatRules.cpp:1410-1535 */
static double onechqs_hn_matel(const double *par, const arr_t *B, const arr_t *S, const mat_t *f, int ist, int jst) {
  double Lambda = par[0], chi = par[1], eN = par[2];
  double Qi = qn_from_st(B, ist, 0), Si = qn_from_st(B, ist, 1);
  double Qj = qn_from_st(B, jst, 0), Sj = qn_from_st(B, jst, 1);
  if (fabs(Qi - Qj) > 1e-20 || fabs(Si - Sj) > 1e-20) return 0.0;
  if (ist == jst) {
    double m = sqrt(Lambda) * B->en.a[ist];
    double Qt = qn_from_st(S, B->itype.a[ist], 0);
    return m + eN * (Qt + 1.0);
  }
  int a, b;
  double c = hn_coupling(B, S, f, ist, jst, 0, &a, &b);
  return chi * get_matel(f, a, b) * c;
}

/* ------------------------------------------------------------------ */
/* BuildBasis  src/BuildBasis.cpp:8-432 (CommonQNs={2,0,1,0,1}, totSpos={1}) */
/* ------------------------------------------------------------------ */
static void filter_qnumbers(arr_t *A) {  /* FilterQNumbers src/NRGarrayClass.cpp:271-309 */
  int nq = A->nq;
  for (int b = 0; b < A->qn.n / nq; b++) {
    for (int c = b + 1; c < A->qn.n / nq;) {
      int eq = 1;
      for (int i = 0; i < nq; i++) if (A->qn.a[c * nq + i] != A->qn.a[b * nq + i]) { eq = 0; break; }
      if (eq) derase(&A->qn, c * nq, (c + 1) * nq); else c++;
    }
  }
}
static void build_basis_qs(const arr_t *Ac, arr_t *Ab, const arr_t *S, int befcut) {
  /* Q adds, S couples (TotalS on position 1); site QN positions {0,1}, Sz at 2.
   * befcut = UpdateBefCut 1: StCameFrom is the parent's pre-cut index (:386-391) */
  Ab->nshell = Ac->nshell + 1;
  Ab->nq = 2;
  arr_clear(Ab);
  for (int b = 0; b < nblocks(Ac); b++) {
    double Qo = getqn(Ac, b, 0), So = getqn(Ac, b, 1);
    for (int s = 0; s < nblocks(S); s++) {
      double Qs = getqn(S, s, 0), Ss = getqn(S, s, 1);
      dpush(&Ab->qn, Qo + Qs); dpush(&Ab->qn, So + Ss);
      for (double sn = So + Ss - 1.0; sn >= fabs(So - Ss); sn -= 1.0) {
        dpush(&Ab->qn, Qo + Qs); dpush(&Ab->qn, sn);
      }
    }
  }
  filter_qnumbers(Ab);
  int istnew = 0;
  for (int bn = 0; bn < nblocks(Ab); bn++) {
    double Qn = getqn(Ab, bn, 0), Sn = getqn(Ab, bn, 1);
    ipush(&Ab->bbe, istnew);
    for (int bo = 0; bo < nblocks(Ac); bo++) {
      double Qo = getqn(Ac, bo, 0), So = getqn(Ac, bo, 1);
      for (int s = 0; s < nblocks(S); s++) {
        double Qs = getqn(S, s, 0), Ss = getqn(S, s, 1), Szs = getqn(S, s, 2);
        int ok = dEqual(Qn, Qs + Qo);
        if (!dEqual(Sn, Szs + So) || !triang(So, Ss, Sn)) ok = 0;
        if (!ok) continue;
        for (int io = blim(Ac, bo, 0); io <= blim(Ac, bo, 1); io++)
          for (int is = blim(S, s, 0); is <= blim(S, s, 1); is++) {
            ipush(&Ab->itype, S->itype.a[is]);
            dpush(&Ab->en, Ac->en.a[io]);
            ipush(&Ab->stcf, befcut == 1 ? Ac->stcf.a[io] : io);
            istnew++;
          }
      }
    }
    ipush(&Ab->bbe, istnew - 1);
  }
}

/* ------------------------------------------------------------------ */
/* DiagHN  src/DiagHN.cpp:6-119                                         */
/* ------------------------------------------------------------------ */
static double diag_hn(const double *par, const arr_t *Ab, const arr_t *S, const mat_t *f, arr_t *Ae,
                      dvec *hrec) {
  Ae->nshell = Ab->nshell; Ae->nq = Ab->nq;
  arr_clear(Ae);
  dcopy_v(&Ae->qn, &Ab->qn); icopy_v(&Ae->bbe, &Ab->bbe);
  if (hrec) hrec->n = 0;
  for (int b = 0; b < nblocks(Ab); b++) {
    int n = bsize(Ab, b), b0 = blim(Ab, b, 0);
    double *ap = (double *)malloc(sizeof(double) * (size_t)n * (size_t)(n + 1) / 2);
    int r = 0;
    for (int i = 0; i < n; i++)
      for (int j = i; j < n; j++) ap[r++] = onechqs_hn_matel(par, Ab, S, f, b0 + i, b0 + j);
    if (hrec) for (int k = 0; k < r; k++) dpush(hrec, ap[k]);
    dres(&Ae->en, Ae->en.n + n);
    dres(&Ae->ev, Ae->ev.n + n * n);
    orc_eigh_packed(n, ap, Ae->en.a + Ae->en.n, Ae->ev.a + Ae->ev.n);
    Ae->en.n += n; Ae->ev.n += n * n;
    free(ap);
  }
  return set_e0zero(Ae);
}

/* ------------------------------------------------------------------ */
/* UpdateMatrices  src/UpdateMatrices.cpp:272-568                       */
/* ------------------------------------------------------------------ */
static void gemm_nt(int m, int n, int k, const double *a, const double *b, double *c) {
  /* c(m x n) = a(m x k) * b(n x k)^T */
  for (int i = 0; i < m; i++)
    for (int j = 0; j < n; j++) {
      double s = 0.0;
      for (int l = 0; l < k; l++) s += a[(size_t)i * k + l] * b[(size_t)j * k + l];
      c[(size_t)i * n + j] = s;
    }
}
static void gemm_nn(int m, int n, int k, const double *a, const double *b, double *c) {
  for (int i = 0; i < m; i++) {
    for (int j = 0; j < n; j++) c[(size_t)i * n + j] = 0.0;
    for (int l = 0; l < k; l++) {
      double x = a[(size_t)i * k + l];
      for (int j = 0; j < n; j++) c[(size_t)i * n + j] += x * b[(size_t)l * n + j];
    }
  }
}
void orc_rotate(int ni, int nj, int ki, int kj, const double *zi, const double *f, const double *zj,
                double *out) {
  double *aux = (double *)malloc(sizeof(double) * (size_t)ni * (size_t)(kj > 0 ? kj : 1));
  gemm_nt(ni, kj, nj, f, zj, aux);
  gemm_nn(ki, kj, ni, zi, aux, out);
  free(aux);
}
static int find_match_block(const arr_t *Ac, int b, const arr_t *Ab) {  /* src/FindMatchBlock.cpp:5-24 */
  double q[8];
  for (int i = 0; i < Ac->nq; i++) q[i] = getqn(Ac, b, i);
  return blk_from_qn(Ab, q);
}
static void update_matrices(const arr_t *S, const arr_t *Ac, const arr_t *Ab, mat_t *mats, int nm) {
  mat_t *old = (mat_t *)calloc((size_t)nm, sizeof(mat_t));
  for (int m = 0; m < nm; m++) {
    if (mats[m].need_old) mat_copy(&old[m], &mats[m]);
    mat_clear(&mats[m]);
    mat_sync(&mats[m], Ac);
  }
  int *icount = (int *)calloc((size_t)nm, sizeof(int));
  for (int ib = 0; ib < nblocks(Ac); ib++) {
    int ni = bsize(Ac, ib);
    int ibc = find_match_block(Ac, ib, Ab);
    int nibc = bsize(Ab, ibc);
    for (int jb = 0; jb < nblocks(Ac); jb++) {
      int nj = bsize(Ac, jb);
      int jbc = find_match_block(Ac, jb, Ab);
      int njbc = bsize(Ab, jbc);
      for (int m = 0; m < nm; m++) {
        if (!mats[m].check(Ac, ib, jb)) continue;
        ipush(&mats[m].map, ib); ipush(&mats[m].map, jb);
        ipush(&mats[m].mbe, icount[m]);
        icount[m] += ni * nj;
        ipush(&mats[m].mbe, icount[m] - 1);
        int zi = blim_ev(Ac, ib), zj = blim_ev(Ac, jb);
        double *F = (double *)malloc(sizeof(double) * (size_t)nibc * (size_t)njbc);
        int r = 0;
        for (int ist = blim(Ab, ibc, 0); ist <= blim(Ab, ibc, 1); ist++)
          for (int jst = blim(Ab, jbc, 0); jst <= blim(Ab, jbc, 1); jst++) {
            double el = mats[m].calc(Ab, S, ist, jst);
            if (mats[m].need_old) el *= get_matel(&old[m], Ab->stcf.a[ist], Ab->stcf.a[jst]);
            F[r++] = el;
          }
        dres(&mats[m].mat, mats[m].mat.n + ni * nj);
        orc_rotate(nibc, njbc, ni, nj, Ac->ev.a + zi, F, Ac->ev.a + zj, mats[m].mat.a + mats[m].mat.n);
        mats[m].mat.n += ni * nj;
        free(F);
      }
    }
  }
  for (int m = 0; m < nm; m++) mat_free(&old[m]);
  free(old); free(icount);
}

/* ------------------------------------------------------------------ */
/* Thermodynamics  src/CalcThermo.cpp:16-121, src/CalcOpAvg.cpp:13-84    */
/* ------------------------------------------------------------------ */
static double calc_suscep(double bb, const arr_t *A) {
  double s0 = 0, s1 = 0;
  for (int b = 0; b < nblocks(A); b++) {
    double S = getqn(A, b, 1);
    double aS = 2.0 * S + 1.0, szsq = (1.0 / 12.0) * aS * (aS * aS - 1.0);
    for (int i = blim(A, b, 0); i <= blim(A, b, 1); i++) {
      double x = exp(-bb * A->en.a[i]);
      s0 += aS * x; s1 += szsq * x;
    }
  }
  return s1 / s0;
}
static double calc_entropy(double bb, const arr_t *A) {
  double Z = 0, s = 0;
  for (int b = 0; b < nblocks(A); b++) {
    double dg = 2.0 * getqn(A, b, 1) + 1.0;
    for (int i = blim(A, b, 0); i <= blim(A, b, 1); i++) {
      double E = A->en.a[i], x = exp(-bb * E);
      s += dg * E * x; Z += dg * x;
    }
  }
  return bb * (s / Z) + log(Z);
}
static double calc_opavg(double bb, const arr_t *A, const mat_t *op) {
  double s = 0, Z = 0;
  for (int b = 0; b < nblocks(A); b++) {
    double dg = 2.0 * getqn(A, b, 1) + 1.0;
    for (int i = blim(A, b, 0); i <= blim(A, b, 1); i++) {
      double x = exp(-bb * A->en.a[i]);
      s += dg * get_matel(op, i, i) * x; Z += dg * x;
    }
  }
  if (fabs(Z) > 1e-20) s /= Z;
  return s;
}

/* ------------------------------------------------------------------ */
/* chain  src/NRGchainClass.cpp:132-155                                 */
/* ------------------------------------------------------------------ */
double orc_chin(int n, double L) {
  double d0 = 1.0 - pow(L, -((double)n + 1.0));
  double d1 = sqrt(1.0 - pow(L, -(2.0 * (double)n + 1.0)));
  double d2 = sqrt(1.0 - pow(L, -(2.0 * (double)n + 3.0)));
  return d0 / (d1 * d2);
}

/* ------------------------------------------------------------------ */
/* NRG_main loop  src/Main/NRG_main.cpp:45-427                          */
/* ------------------------------------------------------------------ */
/* ================================================================== */
/* Derivative of the fixed-truncation NRG_main loop with respect to    */
/* (U, Gamma, ed), forward mode.  No counterpart in the reference.      */
/* Restated from first-order perturbation theory of each symmetric     */
/* block (Z rows = eigenvectors, E ascending):                         */
/*   G = Z dH Z^T,  dE_i = G_ii;                                        */
/*   a kept row moves only towards the discarded rows of its block,    */
/*     dz_i = sum_{j discarded} G_ij / (E_i - E_j) z_j,                 */
/*   so the kept-space Hamiltonian tangent is G on the kept rows and    */
/*   columns (full: rotations among kept states are not undone, every   */
/*   later quantity is invariant under them) and an operator            */
/*   O' = Z_k F Z_k^T moves by dZ_k F Z_k^T + Z_k F dZ_k^T + Z_k dF Z_k^T. */
/*   Thermal sums over a non-diagonal tangent use the divided           */
/*   differences of exp(-beta E) (Daleckii-Krein).                     */
/* The kept set of every step is frozen (given, or the base run's).    */
/* ================================================================== */
typedef struct {
  int ntan;
  double dpar[ORC_MAX_TAN][3];      /* (dU, dGamma, ded) per tangent */
  dvec th[ORC_MAX_TAN];             /* kept-space H tangent: AeigCut block b, k x k at th_off[b] */
  ivec th_off;
  dvec tm[ORC_MAX_TAN][4];          /* operator tangents, element layout of mats[m] */
  dvec G[ORC_MAX_TAN];              /* G of every Aeig block, n x n at ae_off[b] */
  ivec ae_off;
  dvec dz[ORC_MAX_TAN];             /* dZ_k of every AeigCut block, laid out as its ev rows */
  dvec den[ORC_MAX_TAN];            /* d(rescaled E) of Aeig */
  double shift[ORC_MAX_TAN];        /* dEmin (SetE0zero) */
  double dsus[ORC_MAX_TAN][4], dent[ORC_MAX_TAN][4], dops[ORC_MAX_TAN][8];
  orc_tan_cb cb;
  void *user;
} orc_tan;

static void tan_free(orc_tan *T) {
  for (int t = 0; t < ORC_MAX_TAN; t++) {
    free(T->th[t].a); free(T->G[t].a); free(T->dz[t].a); free(T->den[t].a);
    for (int m = 0; m < 4; m++) free(T->tm[t][m].a);
  }
  free(T->th_off.a); free(T->ae_off.a);
}
/* element of an operator tangent stored with the block layout of M */
static double tan_matel(const mat_t *M, const dvec *tv, int ist, int jst) {
  int pi, pj;
  int b1 = blk_from_st(&M->b, ist, &pi), b2 = blk_from_st(&M->b, jst, &pj);
  int im = find_mat_block(M, b1, b2);
  if (im < 0) return 0.0;
  return tv->a[M->mbe.a[2 * im] + bsize(&M->b, b2) * pi + pj];
}
/* dH of one Abasis block (n x n, full): sqrt(Lambda) th(parent, parent') on
 * equal site types, plus chi df + dchi f through the coupling rule */
static void tan_hn_block(const orc_tan *T, int t, double sqrtL, double chi, double dchi, const arr_t *Ab,
                         const arr_t *S, const mat_t *f, const arr_t *Ac, int b0, int n, double *dH) {
  for (int i = 0; i < n; i++)
    for (int j = i; j < n; j++) {
      int ist = b0 + i, jst = b0 + j;
      double v = 0.0;
      if (Ab->itype.a[ist] == Ab->itype.a[jst]) {
        int pp, pq;
        int bp = blk_from_st(Ac, Ab->stcf.a[ist], &pp), bq = blk_from_st(Ac, Ab->stcf.a[jst], &pq);
        if (bp == bq) v += sqrtL * T->th[t].a[T->th_off.a[bp] + pp * bsize(Ac, bp) + pq];
      }
      if (i != j) {
        int a, b;
        double c = hn_coupling(Ab, S, f, ist, jst, 1, &a, &b);
        if (c != 0.0) v += chi * tan_matel(f, &T->tm[t][0], a, b) * c + dchi * get_matel(f, a, b) * c;
      }
      dH[i * n + j] = v;
      dH[j * n + i] = v;
    }
}
/* G = Z dH Z^T for every block of Aeig, the shift dEmin and dE */
static void tan_diag(orc_tan *T, double sqrtL, double chi, const double *dchi, const arr_t *Ab, const arr_t *S,
                     const mat_t *f, const arr_t *Ac, const arr_t *Ae) {
  int nb = nblocks(Ae), ns = Ae->en.n, tot = 0, nmax = 1;
  T->ae_off.n = 0;
  for (int b = 0; b < nb; b++) {
    int n = bsize(Ae, b);
    ipush(&T->ae_off, tot);
    tot += n * n;
    if (n > nmax) nmax = n;
  }
  double *dH = (double *)malloc(sizeof(double) * (size_t)nmax * nmax);
  double *Y = (double *)malloc(sizeof(double) * (size_t)nmax * nmax);
  int gs = 0;   /* SetE0zero's minimum: the first lowest level */
  for (int i = 1; i < ns; i++) if (Ae->en.a[i] < Ae->en.a[gs]) gs = i;
  int gp, gb = blk_from_st(Ae, gs, &gp);
  for (int t = 0; t < T->ntan; t++) {
    dres(&T->G[t], tot); T->G[t].n = tot;
    dres(&T->den[t], ns); T->den[t].n = ns;
    for (int b = 0; b < nb; b++) {
      int n = bsize(Ae, b);
      const double *Z = Ae->ev.a + T->ae_off.a[b];
      tan_hn_block(T, t, sqrtL, chi, dchi[t], Ab, S, f, Ac, blim(Ab, b, 0), n, dH);
      gemm_nn(n, n, n, Z, dH, Y);
      gemm_nt(n, n, n, Y, Z, T->G[t].a + T->ae_off.a[b]);
    }
    double s = T->G[t].a[T->ae_off.a[gb] + gp * bsize(Ae, gb) + gp];
    T->shift[t] = s;
    for (int b = 0; b < nb; b++) {
      int n = bsize(Ae, b), b0 = blim(Ae, b, 0);
      for (int i = 0; i < n; i++) T->den[t].a[b0 + i] = T->G[t].a[T->ae_off.a[b] + i * n + i] - s;
    }
  }
  free(dH); free(Y);
}
/* CalcSuscep / CalcEntropy (src/CalcThermo.cpp:16-121) and their derivatives along dE */
static void tan_thermo(double bb, const arr_t *A, const double *dE, double *dsus, double *dent) {
  double s0 = 0, s1 = 0, d0 = 0, d1 = 0, Z = 0, se = 0, dZ = 0, dse = 0;
  for (int b = 0; b < nblocks(A); b++) {
    double S = getqn(A, b, 1);
    double aS = 2.0 * S + 1.0, szsq = (1.0 / 12.0) * aS * (aS * aS - 1.0);
    for (int i = blim(A, b, 0); i <= blim(A, b, 1); i++) {
      double E = A->en.a[i], x = exp(-bb * E), dx = -bb * dE[i] * x;
      s0 += aS * x; s1 += szsq * x; d0 += aS * dx; d1 += szsq * dx;
      Z += aS * x; dZ += aS * dx; se += aS * E * x; dse += aS * (dE[i] * x + E * dx);
    }
  }
  *dsus = (d1 * s0 - s1 * d0) / (s0 * s0);
  *dent = bb * (dse * Z - se * dZ) / (Z * Z) + dZ / Z;
}
/* the new kept-space tangent, C = G_kd / (E_k - E_d) and dZ_k = C Z_d per AeigCut block */
static void tan_cut(orc_tan *T, const arr_t *Ae, const arr_t *Ac) {
  int nbc = nblocks(Ac), tot = 0;
  T->th_off.n = 0;
  for (int b = 0; b < nbc; b++) { ipush(&T->th_off, tot); tot += bsize(Ac, b) * bsize(Ac, b); }
  for (int t = 0; t < T->ntan; t++) {
    dres(&T->th[t], tot); T->th[t].n = tot;
    dres(&T->dz[t], Ac->ev.n); T->dz[t].n = Ac->ev.n;
    for (int b = 0; b < nbc; b++) {
      int k = bsize(Ac, b), n = bsize_bc(Ac, b), be = 0;
      while (blim(Ae, be, 0) != Ac->bbe_bc.a[2 * b]) be++;
      const double *G = T->G[t].a + T->ae_off.a[be];
      const double *E = Ae->en.a + blim(Ae, be, 0);
      const double *Z = Ae->ev.a + T->ae_off.a[be];
      double *th = T->th[t].a + T->th_off.a[b];
      for (int i = 0; i < k; i++)
        for (int j = 0; j < k; j++) th[i * k + j] = G[i * n + j] - (i == j ? T->shift[t] : 0.0);
      double *dz = T->dz[t].a + blim_ev(Ac, b);
      if (k == n) { memset(dz, 0, sizeof(double) * (size_t)k * n); continue; }
      double *C = (double *)malloc(sizeof(double) * (size_t)k * (n - k));
      for (int i = 0; i < k; i++)
        for (int j = k; j < n; j++) {
          double d = E[i] - E[j];
          C[i * (n - k) + (j - k)] = d != 0.0 ? G[i * n + j] / d : 0.0;
        }
      gemm_nn(k, n, n - k, C, Z + (size_t)k * n, dz);
      free(C);
    }
  }
}
/* operator tangents over the new AeigCut, in UpdateMatrices' (ib, jb, m) order;
 * called with the operators of the previous step still in mats */
static void tan_ops(orc_tan *T, const arr_t *S, const arr_t *Ac, const arr_t *Ab, const mat_t *mats, int nm) {
  dvec nt[ORC_MAX_TAN][4];
  memset(nt, 0, sizeof(nt));
  for (int ib = 0; ib < nblocks(Ac); ib++) {
    int ki = bsize(Ac, ib), ibc = find_match_block(Ac, ib, Ab), ni = bsize(Ab, ibc);
    for (int jb = 0; jb < nblocks(Ac); jb++) {
      int kj = bsize(Ac, jb), jbc = find_match_block(Ac, jb, Ab), nj = bsize(Ab, jbc);
      for (int m = 0; m < nm; m++) {
        if (!mats[m].check(Ac, ib, jb)) continue;
        double *F = (double *)malloc(sizeof(double) * (size_t)ni * nj);
        double *dF = (double *)malloc(sizeof(double) * (size_t)ni * nj);
        double *o = (double *)malloc(sizeof(double) * (size_t)ki * kj);
        const double *Zi = Ac->ev.a + blim_ev(Ac, ib), *Zj = Ac->ev.a + blim_ev(Ac, jb);
        int r = 0;
        for (int ist = blim(Ab, ibc, 0); ist <= blim(Ab, ibc, 1); ist++)
          for (int jst = blim(Ab, jbc, 0); jst <= blim(Ab, jbc, 1); jst++, r++) {
            double g = mats[m].calc(Ab, S, ist, jst);
            F[r] = mats[m].need_old ? g * get_matel(&mats[m], Ab->stcf.a[ist], Ab->stcf.a[jst]) : g;
          }
        for (int t = 0; t < T->ntan; t++) {
          const double *dZi = T->dz[t].a + blim_ev(Ac, ib), *dZj = T->dz[t].a + blim_ev(Ac, jb);
          dres(&nt[t][m], nt[t][m].n + ki * kj);
          double *out = nt[t][m].a + nt[t][m].n;
          orc_rotate(ni, nj, ki, kj, dZi, F, Zj, out);
          orc_rotate(ni, nj, ki, kj, Zi, F, dZj, o);
          for (int q = 0; q < ki * kj; q++) out[q] += o[q];
          if (mats[m].need_old) {
            r = 0;
            for (int ist = blim(Ab, ibc, 0); ist <= blim(Ab, ibc, 1); ist++)
              for (int jst = blim(Ab, jbc, 0); jst <= blim(Ab, jbc, 1); jst++, r++)
                dF[r] = mats[m].calc(Ab, S, ist, jst) *
                        tan_matel(&mats[m], &T->tm[t][m], Ab->stcf.a[ist], Ab->stcf.a[jst]);
            orc_rotate(ni, nj, ki, kj, Zi, dF, Zj, o);
            for (int q = 0; q < ki * kj; q++) out[q] += o[q];
          }
          nt[t][m].n += ki * kj;
        }
        free(F); free(dF); free(o);
      }
    }
  }
  for (int t = 0; t < T->ntan; t++)
    for (int m = 0; m < nm; m++) { free(T->tm[t][m].a); T->tm[t][m] = nt[t][m]; }
}
/* derivative of CalcOpAvg (src/CalcOpAvg.cpp:13-84) on AeigCut: the kept-space
 * Hamiltonian is diag(E) + th, so sum_i O_ii x_i becomes Tr[O exp(-beta H)] */
static double tan_opavg(double bb, const arr_t *A, const mat_t *op, const dvec *top, const orc_tan *T, int t) {
  double s = 0, Z = 0, ds = 0, dZ = 0;
  for (int b = 0; b < nblocks(A); b++) {
    double dg = 2.0 * getqn(A, b, 1) + 1.0;
    int k = bsize(A, b), b0 = blim(A, b, 0);
    int im = find_mat_block(op, b, b);
    const double *O = im >= 0 ? op->mat.a + op->mbe.a[2 * im] : NULL;
    const double *dO = im >= 0 ? top->a + op->mbe.a[2 * im] : NULL;
    const double *th = T->th[t].a + T->th_off.a[b];
    const double *E = A->en.a + b0;
    for (int i = 0; i < k; i++) {
      double x = exp(-bb * E[i]);
      Z += dg * x;
      dZ += dg * th[i * k + i] * (-bb * x);
      if (!O) continue;
      s += dg * O[i * k + i] * x;
      ds += dg * dO[i * k + i] * x;
      for (int j = 0; j < k; j++) {
        double d = E[i] - E[j], xj = exp(-bb * E[j]);
        double phi = d != 0.0 ? xj * expm1(-bb * d) / d : -bb * xj;
        ds += dg * O[j * k + i] * th[i * k + j] * phi;
      }
    }
  }
  return (ds * Z - s * dZ) / (Z * Z);
}
/* CutStates with a given Kept set (src/CutStates.cpp:5-96 after SetKept):
 * blocks without kept states go, then each block keeps its kept prefix */
static void cut_states_fixed(const arr_t *Ain, arr_t *Ac) {
  arr_clear(Ac);
  Ac->nshell = Ain->nshell; Ac->nq = Ain->nq;
  dcopy_v(&Ac->qn, &Ain->qn); dcopy_v(&Ac->en, &Ain->en); dcopy_v(&Ac->ev, &Ain->ev);
  icopy_v(&Ac->kept, &Ain->kept); icopy_v(&Ac->bbe, &Ain->bbe);
  Ac->stcf.n = 0;
  for (int i = 0; i < Ac->en.n; i++) ipush(&Ac->stcf, i);
  icopy_v(&Ac->bbe_bc, &Ain->bbe);
  set_bbe_ev(Ac);
  for (int b = 0; b < nblocks(Ac); b++) {
    int nk = 0;
    for (int i = blim(Ac, b, 0); i <= blim(Ac, b, 1); i++) nk += Ac->kept.a[i] != 0;
    if (nk == 0) { remove_block(Ac, b); b--; }
  }
  for (int b = 0; b < nblocks(Ac); b++) {
    int b0 = blim(Ac, b, 0), b1 = blim(Ac, b, 1), nk = 0;
    for (int i = b0; i <= b1; i++) nk += Ac->kept.a[i] != 0;
    if (b0 + nk <= b1) remove_states_from_block(Ac, b, b0 + nk, b1);
  }
}

static void emit_tan(orc_step_cb cb, void *user, const arr_t *Ae, const arr_t *Ac, double emin, double ecut,
                     int nkept, const dvec *hrec, int rec, const double *thT, const double *sus,
                     const double *ent, double Tavg, int nops, const double *ops, const void *tan);
static void emit(orc_step_cb cb, void *user, const arr_t *Ae, const arr_t *Ac, double emin, double ecut,
                 int nkept, const dvec *hrec, int rec, const double *thT, const double *sus,
                 const double *ent, double Tavg, int nops, const double *ops) {
  emit_tan(cb, user, Ae, Ac, emin, ecut, nkept, hrec, rec, thT, sus, ent, Tavg, nops, ops, NULL);
}
static void emit_tan(orc_step_cb cb, void *user, const arr_t *Ae, const arr_t *Ac, double emin, double ecut,
                     int nkept, const dvec *hrec, int rec, const double *thT, const double *sus,
                     const double *ent, double Tavg, int nops, const double *ops, const void *tan) {
  if (!cb && !tan) return;
  orc_step st;
  memset(&st, 0, sizeof(st));
  st.nshell = Ae->nshell; st.nq = Ae->nq; st.nblocks = nblocks(Ae);
  st.qn = Ae->qn.a; st.bbe = Ae->bbe.a; st.nstates = Ae->en.n; st.en = Ae->en.a;
  unsigned char *kp = (unsigned char *)malloc((size_t)(Ae->en.n + 1));
  for (int i = 0; i < Ae->en.n; i++) kp[i] = (i < Ae->kept.n) ? (unsigned char)Ae->kept.a[i] : 1;
  st.kept = kp;
  st.emin = emin; st.ecut = ecut; st.nkept = nkept;
  st.hpacked = (rec && hrec) ? hrec->a : NULL;
  st.evec = rec ? Ae->ev.a : NULL;
  for (int i = 0; i < 4; i++) { st.thermo_T[i] = thT ? thT[i] : 0; st.suscep[i] = sus ? sus[i] : 0; st.entropy[i] = ent ? ent[i] : 0; }
  st.T_avg = Tavg; st.nops_avg = nops;
  for (int i = 0; i < nops && i < 8; i++) st.opavg[i] = ops[i];
  st.cut_nblocks = nblocks(Ac); st.cut_qn = Ac->qn.a; st.cut_bbe = Ac->bbe.a; st.cut_bbe_bc = Ac->bbe_bc.a;
  if (tan) {
    const orc_tan *T = (const orc_tan *)tan;
    orc_tan_step dt;
    memset(&dt, 0, sizeof(dt));
    dt.nshell = Ae->nshell; dt.ntan = T->ntan; dt.nstates = Ae->en.n; dt.nops_avg = nops;
    double *den = (double *)malloc(sizeof(double) * (size_t)(T->ntan * Ae->en.n + 1));
    for (int t = 0; t < T->ntan; t++) {
      memcpy(den + (size_t)t * Ae->en.n, T->den[t].a, sizeof(double) * (size_t)Ae->en.n);
      for (int i = 0; i < 4; i++) { dt.dsuscep[t][i] = T->dsus[t][i]; dt.dentropy[t][i] = T->dent[t][i]; }
      for (int i = 0; i < nops && i < 8; i++) dt.dopavg[t][i] = T->dops[t][i];
    }
    dt.den = den;
    if (T->cb) T->cb(T->user, &st, &dt);
    free(den);
  } else {
    cb(user, &st);
  }
  free(kp);
}

struct orc_dm;
static void dm_store_step(struct orc_dm *D, const arr_t *Ac, const arr_t *Ab, const mat_t *op);
static void dm_store_tan(struct orc_dm *D, const orc_tan *T);
static int run_onechqs(const orc_onechqs_params *p, orc_step_cb cb, void *user, struct orc_dm *dm, int befcut,
                       const unsigned char *kfix, const long long *kfix_off, orc_tan *T);
int orc_run_onechqs(const orc_onechqs_params *p, orc_step_cb cb, void *user) {
  return run_onechqs(p, cb, user, NULL, 0, NULL, NULL, NULL);
}
/* The NRG_main loop with every step's Kept set given (kept + kept_off[N] for
 * N = 0 .. nsitesmax-1, Aeig state order; NULL: SetKept as usual) and ntan
 * tangents along dpar[t] = (dU, dGamma, ded). */
int orc_run_onechqs_jvp(const orc_onechqs_params *p, const unsigned char *kept, const long long *kept_off, int ntan,
                        const double *dpar, orc_tan_cb cb, void *user) {
  if (ntan < 0 || ntan > ORC_MAX_TAN || (ntan > 0 && !dpar) || (kept && !kept_off)) return -1;
  if (ntan > 0 && !(p->Gamma > 0.0)) return -1;
  orc_tan T;
  memset(&T, 0, sizeof(T));
  T.ntan = ntan; T.cb = cb; T.user = user;
  for (int t = 0; t < ntan; t++)
    for (int q = 0; q < 3; q++) T.dpar[t][q] = dpar[3 * t + q];
  int rc = run_onechqs(p, NULL, NULL, NULL, 0, kept, kept_off, &T);
  tan_free(&T);
  return rc;
}
/* befcut = UpdateBefCut 1 (src/Main/NRG_main.cpp:334-343,393-396): the operators are
 * rotated over every eigenstate of AeigCut = Aeig with its Kept flags, SaveArrays
 * writes that copy, and the real CutStates follows CalcStuff */
static int run_onechqs(const orc_onechqs_params *p, orc_step_cb cb, void *user, struct orc_dm *dm, int befcut,
                       const unsigned char *kfix, const long long *kfix_off, orc_tan *T) {
  double Lambda = p->Lambda, bbar = p->betabar > 0 ? p->betabar : 0.727;
  double myeps = p->myeps > 0 ? p->myeps : 1e-18;
  double HLF = 0.5 * (1.0 + 1.0 / Lambda);
  arr_t S = {0}, Ae = {0}, Ac = {0}, Ab = {0};
  onechqs_site(&S);
  /* operators: [0] f_N; calcdens 0/2 -> [1..3] n_d, n_d^2, S_d^2; calcdens 1 -> [1] c_d */
  int nm = (p->calcdens == 1) ? 2 : 4;
  mat_t mats[4];
  memset(mats, 0, sizeof(mats));
  mats[0].need_old = 0; mats[0].upper = 0; mats[0].check = onechqs_cd_check; mats[0].calc = onechqs_fn_matel;
  for (int m = 1; m < nm; m++) {
    mats[m].need_old = 1; mats[m].upper = 0;
    if (p->calcdens == 1) { mats[m].check = onechqs_cd_check; mats[m].calc = onechqs_cd_matel; }
    else { mats[m].check = diag_check; mats[m].calc = imponly_matel; mats[m].calc_avg = 1; }
  }
  /* OneChQS_SetAnderson_Hm1  src/Main/Main_SetH0.cpp:1794-1969 */
  double chi_m1 = sqrt(2.0 * p->Gamma / M_PI) / (sqrt(Lambda) * HLF);
  if (p->disc_scheme != 1) chi_m1 *= sqrt(0.5 * log(Lambda) * (Lambda + 1) / (Lambda - 1));
  Ae.nq = 2; Ae.nshell = -1;
  const double q0[6] = {-1, 0, 0, 0.5, 1, 0};
  for (int i = 0; i < 6; i++) dpush(&Ae.qn, q0[i]);
  for (int i = 0; i < 3; i++) { ipush(&Ae.bbe, i); ipush(&Ae.bbe, i); dpush(&Ae.ev, 1.0); }
  dpush(&Ae.en, 0.5 * p->U / (Lambda * HLF));
  dpush(&Ae.en, (p->ed + 0.5 * p->U) / (Lambda * HLF));
  dpush(&Ae.en, (2.0 * p->ed + 1.5 * p->U) / (Lambda * HLF));
  for (int m = 0; m < nm; m++) { mats[m].b.nshell = -1; mat_sync(&mats[m], &Ae); }
  dpush(&mats[0].mat, 1.0); ipush(&mats[0].map, 0); ipush(&mats[0].map, 1); ipush(&mats[0].mbe, 0); ipush(&mats[0].mbe, 0);
  dpush(&mats[0].mat, -sqrt(2.0)); ipush(&mats[0].map, 1); ipush(&mats[0].map, 2); ipush(&mats[0].mbe, 1); ipush(&mats[0].mbe, 1);
  if (p->calcdens == 1) {
    mat_t tmp = {0}; mat_copy(&tmp, &mats[0]);
    tmp.need_old = 1; tmp.check = onechqs_cd_check; tmp.calc = onechqs_cd_matel;
    mat_copy(&mats[1], &tmp); mat_free(&tmp);
  } else {
    for (int m = 1; m < nm; m++) mat_set_diag(&mats[m], 0.0);
    for (int b = 0; b < 3; b++) {
      double nd = getqn(&Ae, b, 0) + 1.0, Sd = getqn(&Ae, b, 1);
      mat_put(&mats[1], b, b, 0, 0, nd);
      mat_put(&mats[2], b, b, 0, 0, nd * nd);
      mat_put(&mats[3], b, b, 0, 0, Sd * (Sd + 1.0));
    }
  }
  cut_states(&Ae, &Ac, p->ncut, myeps);
  double DN = HLF * pow(Lambda, -(0 - 2) / 2.0);
  if (T) {   /* tangents of H_{-1}: one-state blocks, f and the impurity operators are constants */
    const double sc = 1.0 / (Lambda * HLF);
    T->th_off.n = 0;
    for (int b = 0; b < 3; b++) ipush(&T->th_off, b);
    for (int t = 0; t < T->ntan; t++) {
      double dU = T->dpar[t][0], ded = T->dpar[t][2];
      T->th[t].n = 0; T->den[t].n = 0;
      double e3[3] = {0.5 * dU * sc, (ded + 0.5 * dU) * sc, (2.0 * ded + 1.5 * dU) * sc};
      for (int b = 0; b < 3; b++) { dpush(&T->th[t], e3[b]); dpush(&T->den[t], e3[b]); }
      for (int m = 0; m < nm; m++) { T->tm[t][m].n = 0; for (int q = 0; q < mats[m].mat.n; q++) dpush(&T->tm[t][m], 0.0); }
    }
  }
  {
    double ops[3] = {0, 0, 0};
    int nops = 0;
    if (p->calcdens != 1) for (int m = 1; m < nm; m++) ops[nops++] = calc_opavg(bbar, &Ac, &mats[m]);
    if (T)
      for (int t = 0; t < T->ntan; t++)
        for (int m = 1, q = 0; m < nm && p->calcdens != 1; m++, q++) T->dops[t][q] = tan_opavg(bbar, &Ac, &mats[m], &T->tm[t][m], T, t);
    emit_tan(cb, user, &Ae, &Ac, 0.0, HUGE_VAL, Ae.en.n, NULL, 0, NULL, NULL, NULL, DN / bbar, nops, ops, T);
  }
  double par[4] = {Lambda, chi_m1, 0.0, 0.0};
  double dchi[ORC_MAX_TAN] = {0, 0, 0, 0};
  if (T) for (int t = 0; t < T->ntan; t++) dchi[t] = chi_m1 * T->dpar[t][1] / (2.0 * p->Gamma);
  dvec hrec = {0};
  for (int N = 0; N < p->nsitesmax; N++) {
    DN = HLF * pow(Lambda, -(N - 1) / 2.0);
    build_basis_qs(&Ac, &Ab, &S, befcut);
    double emin = diag_hn(par, &Ab, &S, &mats[0], &Ae, p->record_blocks ? &hrec : NULL);
    if (T) tan_diag(T, sqrt(Lambda), par[1], dchi, &Ab, &S, &mats[0], &Ac, &Ae);
    int nkept = 0;
    double ecut;
    if (kfix) {   /* the given Kept set; Ecut reported as its highest level */
      orc_setdegen(Ae.en.n, Ae.en.a, myeps);
      Ae.kept.n = 0;
      ecut = 0.0;
      for (int i = 0; i < Ae.en.n; i++) {
        int k = kfix[kfix_off[N] + i] != 0;
        ipush(&Ae.kept, k); nkept += k;
        if (k && Ae.en.a[i] > ecut) ecut = Ae.en.a[i];
      }
    } else {
      ecut = set_kept(&Ae, p->ncut, myeps, &nkept);
    }
    double thT[4], sus[4], ent[4];
    for (int nc = 1; nc <= 4; nc++) {
      double bN = bbar * pow(Lambda, (nc - 4) / (2.0 * 4));
      thT[nc - 1] = DN / bN;
      sus[nc - 1] = calc_suscep(bN, &Ae);
      ent[nc - 1] = calc_entropy(bN, &Ae);
      if (T)
        for (int t = 0; t < T->ntan; t++) tan_thermo(bN, &Ae, T->den[t].a, &T->dsus[t][nc - 1], &T->dent[t][nc - 1]);
    }
    if (kfix) cut_states_fixed(&Ae, &Ac);
    else cut_states(&Ae, &Ac, befcut ? 2 * Ae.en.n : p->ncut, myeps);
    if (T) { tan_cut(T, &Ae, &Ac); tan_ops(T, &S, &Ac, &Ab, mats, nm); }
    update_matrices(&S, &Ac, &Ab, mats, nm);
    if (dm && p->calcdens == 1) dm_store_step(dm, &Ac, &Ab, &mats[1]);   /* SaveArrays src/NRGcodehandlerClass.cpp:887-933 */
    if (dm && T && p->calcdens == 1) dm_store_tan(dm, T);
    double ops[3] = {0, 0, 0};
    int nops = 0;
    if (p->calcdens != 1) for (int m = 1; m < nm; m++) ops[nops++] = calc_opavg(bbar, &Ac, &mats[m]);
    if (T)
      for (int t = 0; t < T->ntan; t++)
        for (int m = 1, q = 0; m < nm && p->calcdens != 1; m++, q++) T->dops[t][q] = tan_opavg(bbar, &Ac, &mats[m], &T->tm[t][m], T, t);
    if (befcut) cut_states(&Ae, &Ac, p->ncut, myeps);
    emit_tan(cb, user, &Ae, &Ac, emin, ecut, nkept, &hrec, p->record_blocks, thT, sus, ent, DN / bbar, nops, ops, T);
    par[1] = orc_chin(N, Lambda);
    for (int t = 0; t < ORC_MAX_TAN; t++) dchi[t] = 0.0;
  }
  free(hrec.a);
  for (int m = 0; m < 4; m++) mat_free(&mats[m]);
  arr_free(&S); arr_free(&Ae); arr_free(&Ac); arr_free(&Ab);
  return 0;
}

/* ================================================================== */
/* TwoChQS Anderson: legacy standalone driver src/TwoChQS/TwoChQS.cpp  */
/* (ModelNo 0) with QS_BuildBasis, TwoChQS_DiagHN,                     */
/* TwoChQS_SetH0Anderson, TwoChQS_UpdateQm1fQ /                        */
/* TwoChQS_UpdateMatrixAfterCutting and the by-value CutStates.        */
/* ================================================================== */

/* TwoChQS_SetSingleSite  src/TwoChQS/TwoChQS_SetSingleSite.cpp: 16 states, 10 (Q,S,Sz) blocks */
static void twochqs_site(arr_t *S) {
  arr_clear(S);
  S->nq = 3; S->nshell = 0;
  const double q[30] = {-2, 0, 0, -1, 0.5, 0.5, -1, 0.5, -0.5, 0, 0, 0, 0, 1, -1,
                        0, 1, 0, 0, 1, 1, 1, 0.5, 0.5, 1, 0.5, -0.5, 2, 0, 0};
  const int be[20] = {0, 0, 1, 2, 3, 4, 5, 7, 8, 8, 9, 9, 10, 10, 11, 12, 13, 14, 15, 15};
  for (int i = 0; i < 30; i++) dpush(&S->qn, q[i]);
  for (int i = 0; i < 20; i++) ipush(&S->bbe, be[i]);
  for (int i = 0; i < 16; i++) ipush(&S->itype, i);
}
/* TwoChQS_fd_table  src/TwoChQS/TwoChQS_OpTables.cpp:19-146 */
static double twoch_fd(int ch, int sigma, int ti, int tj) {
  static double M[2][2][16][16];
  static int init = 0;
  if (!init) {
    const double r = 1.0 / sqrt(2.0);
    memset(M, 0, sizeof(M));
    double (*a)[16] = M[0][1]; /* ch1 up */
    a[2][0] = 1; a[10][1] = 1; a[9][3] = r; a[11][6] = 1; a[5][3] = r; a[7][4] = 1; a[12][9] = r;
    a[12][5] = -r; a[14][8] = 1; a[15][13] = 1;
    a = M[0][0]; /* ch1 dn */
    a[4][0] = 1; a[9][1] = r; a[5][1] = -r; a[8][3] = 1; a[13][6] = 1; a[7][2] = -1; a[12][10] = -1;
    a[14][9] = -r; a[14][5] = -r; a[15][11] = -1;
    a = M[1][1]; /* ch2 up */
    a[1][0] = 1; a[6][3] = 1; a[10][2] = -1; a[11][9] = -r; a[11][5] = -r; a[9][4] = -r; a[5][4] = r;
    a[13][8] = -1; a[12][7] = 1; a[15][14] = 1;
    a = M[1][0]; /* ch2 dn */
    a[3][0] = 1; a[6][1] = -1; a[9][2] = -r; a[11][10] = 1; a[5][2] = -r; a[8][4] = -1; a[13][9] = r;
    a[13][5] = -r; a[14][7] = 1; a[15][12] = -1;
    init = 1;
  }
  if (ch < 1 || ch > 2 || (sigma != 1 && sigma != -1)) return 0.0;
  if (ti < 0 || ti > 15 || tj < 0 || tj > 15) return 0.0;
  return M[ch - 1][sigma == 1 ? 1 : 0][ti][tj];
}
/* TwoChQS_SpSm_table  src/TwoChQS/TwoChQS_OpTables.cpp:148-183 */
static double twoch_spsm(int ti, int tj) {
  if (ti == tj) return 1.0;
  int a = ti < tj ? ti : tj, b = ti < tj ? tj : ti;
  if ((a == 1 && b == 3) || (a == 2 && b == 4) || (a == 11 && b == 13) || (a == 12 && b == 14) ||
      (a == 8 && b == 9) || (a == 9 && b == 10) || (a == 8 && b == 10))
    return 1.0;
  return 0.0;
}
static int dLEqual(double a, double b) { return (a < b) || fabs(a - b) < 1e-20; }

/* QS_BuildBasis  src/QS_BuildBasis.cpp:9-169 (KeepSz=0): block list as in
 * BuildBasis; states ordered (old block, site block, site state, old state). */
static void qs_build_basis(const arr_t *Ac, arr_t *Ab, const arr_t *S, int befcut) {
  Ab->nshell = Ac->nshell + 1;
  Ab->nq = Ac->nq;
  arr_clear(Ab);
  for (int b = 0; b < nblocks(Ac); b++) {
    double Qo = getqn(Ac, b, 0), So = getqn(Ac, b, 1);
    for (int s = 0; s < nblocks(S); s++) {
      double Qs = getqn(S, s, 0), Ss = getqn(S, s, 1);
      for (double sn = So + Ss; sn >= fabs(So - Ss); sn -= 1.0) { dpush(&Ab->qn, Qo + Qs); dpush(&Ab->qn, sn); }
    }
  }
  filter_qnumbers(Ab);
  int istnew = 0;
  for (int bn = 0; bn < nblocks(Ab); bn++) {
    ipush(&Ab->bbe, istnew);
    double Qn = getqn(Ab, bn, 0), Sn = getqn(Ab, bn, 1);
    for (int bo = 0; bo < nblocks(Ac); bo++) {
      double Qo = getqn(Ac, bo, 0), So = getqn(Ac, bo, 1);
      for (int s = 0; s < nblocks(S); s++) {
        double Qs = getqn(S, s, 0), Ss = getqn(S, s, 1), Szs = getqn(S, s, 2);
        for (int is = blim(S, s, 0); is <= blim(S, s, 1); is++) {
          if (!(dEqual(Qn, Qo + Qs) && dEqual(Sn, So + Szs) && triang(So, Ss, Sn))) continue;
          for (int io = blim(Ac, bo, 0); io <= blim(Ac, bo, 1); io++) {
            ipush(&Ab->itype, S->itype.a[is]);
            dpush(&Ab->en, Ac->en.a[io]);
            ipush(&Ab->stcf, befcut == 1 ? Ac->stcf.a[io] : io);
            istnew++;
          }
        }
      }
    }
    ipush(&Ab->bbe, istnew - 1);
  }
}

/* site state range of a site block; block -1 gives [0,0] as GetBlockLimit(-1,.) returns 0 */
static void site_range(const arr_t *S, int blk, int *b0, int *b1) {
  if (blk < 0) { *b0 = 0; *b1 = 0; return; }
  *b0 = S->bbe.a[2 * blk]; *b1 = S->bbe.a[2 * blk + 1];
}

/* coefficient of chi*<old|f_ch|old'> in an H_N element: the sigma/Szold sums of
 * TwoChQS_DiagHN (src/TwoChQS/TwoChQS_DiagHN.cpp:62-250).  h0 selects the
 * TwoChQS_SetH0Anderson variant of the fermion sign (Q and Sz test). */
static double twoch_hn_coef(const arr_t *S, int ch, double Si, int typep, int type, int h0) {
  int pos;
  int bp = blk_from_st(S, typep, &pos), bb = blk_from_st(S, type, &pos);
  double Qtp = getqn(S, bp, 0), Stp = getqn(S, bp, 1), Sztp = getqn(S, bp, 2);
  double Qt = getqn(S, bb, 0), St = getqn(S, bb, 1), Szt = getqn(S, bb, 2);
  double Soldp = Si - Sztp, Sold = Si - Szt;
  double fs = 1.0;
  if (h0) { if ((dEqual(Qt, 1.0) || dEqual(Qt, -1.0)) && (dEqual(Szt, 0.5) || dEqual(Szt, -0.5))) fs = -1.0; }
  else if (dEqual(Szt, 0.5) || dEqual(Szt, -0.5)) fs = -1.0;
  double qs[3] = {Qt, St, Szt}, qsp[3] = {Qtp, Stp, Sztp};
  double tot = 0.0;
  for (int sg = -1; sg <= 1; sg += 2) {
    double ds = 0.5 * sg, full = 0.0;
    for (double szo = Sold; szo >= -Sold; szo -= 1.0) {
      double szop = szo - ds, szt = Si - szo, sztp = szt + ds;
      qs[2] = szt; qsp[2] = sztp;
      int sb = blk_from_qn(S, qs), sbp = blk_from_qn(S, qsp);
      if (dLEqual(fabs(sztp), Stp) && dLEqual(fabs(szo), Sold) && dLEqual(fabs(szop), Soldp)) {
        double c0 = orc_cgordan(Sold, szo, St, szt, Si, Si);
        double c1 = orc_cgordan(Soldp, szop, Stp, sztp, Si, Si);
        double c2 = orc_cgordan(Soldp, szop, 0.5, ds, Sold, szo);
        int p0, p1, q0, q1;
        site_range(S, sbp, &p0, &p1);
        site_range(S, sb, &q0, &q1);
        for (int ssp = p0; ssp <= p1; ssp++)
          for (int ss = q0; ss <= q1; ss++)
            full += c0 * c1 * c2 * twoch_fd(ch, sg, ssp, ss) * fs * twoch_spsm(ss, type) * twoch_spsm(ssp, typep);
      }
    }
    tot += full;
  }
  return tot;
}

/* element of f_ch between basis states with the same parent: the fnbasis rule of
 * TwoChQS_UpdateQm1fQ / TwoChQS_UpdateMatrixAfterCutting
 * (src/TwoChQS/TwoChQS_UpdateMatrices.cpp:90-196, 380-470) */
static double twoch_fn_el(const arr_t *S, int ch, double Si, double Sj, int ti, int tj) {
  int pos;
  int bi = blk_from_st(S, ti, &pos), bj = blk_from_st(S, tj, &pos);
  double Qti = getqn(S, bi, 0), Sti = getqn(S, bi, 1), Szti = getqn(S, bi, 2);
  double Qtj = getqn(S, bj, 0), Stj = getqn(S, bj, 1), Sztj = getqn(S, bj, 2);
  double qi[3] = {Qti, Sti, Szti}, qj[3] = {Qtj, Stj, Sztj};
  double Scfi = Si - Szti, Scfj = Sj - Sztj;
  double cgm = orc_cgordan(Si, Si, 0.5, -0.5, Sj, Sj), cgp = orc_cgordan(Si, Si, 0.5, 0.5, Sj, Sj);
  int sigma = 0;
  double norm = 1.0;
  if (fabs(cgm) > 1e-20) { sigma = -1; norm = cgm; }
  else if (fabs(cgp) > 1e-20) { sigma = 1; norm = cgp; }
  if (sigma == 0) return 0.0;
  double acc = 0.0;
  for (double szc = Scfj; szc >= -Scfj; szc -= 1.0) {
    double szi = Si - szc, szj = Sj - szc;
    qi[2] = szi; qj[2] = szj;
    int bsi = blk_from_qn(S, qi), bsj = blk_from_qn(S, qj);
    double c0 = orc_cgordan(Scfi, szc, Sti, szi, Si, Si);
    double c1 = orc_cgordan(Scfj, szc, Stj, szj, Sj, Sj);
    int a0, a1, b0, b1;
    site_range(S, bsi, &a0, &a1);
    site_range(S, bsj, &b0, &b1);
    for (int ssi = a0; ssi <= a1; ssi++)
      for (int ssj = b0; ssj <= b1; ssj++)
        acc += twoch_fd(ch, sigma, ssj, ssi) * c0 * c1 * twoch_spsm(ssi, ti) * twoch_spsm(ssj, tj);
  }
  return acc / norm;
}

/* by-value CutStates(CNRGarray*,int)  src/CutStates.cpp:103-160: no SetKept/SetDegen on the copy */
static void cut_states_legacy(const arr_t *Ain, arr_t *Ac, int ncut) {
  arr_clear(Ac);
  Ac->nshell = Ain->nshell; Ac->nq = Ain->nq;
  dcopy_v(&Ac->qn, &Ain->qn); dcopy_v(&Ac->en, &Ain->en); dcopy_v(&Ac->ev, &Ain->ev);
  icopy_v(&Ac->bbe, &Ain->bbe);
  for (int i = 0; i < Ac->en.n; i++) { ipush(&Ac->stcf, i); ipush(&Ac->kept, 1); }
  icopy_v(&Ac->bbe_bc, &Ain->bbe);
  set_bbe_ev(Ac);
  if (ncut >= nstates(Ain)) return;
  double ec = orc_ecut(Ain->en.n, Ain->en.a, ncut);
  for (int b = 0; b < nblocks(Ac); b++) {
    int b0 = blim(Ac, b, 0), b1 = blim(Ac, b, 1);
    double emin = Ac->en.a[b0];
    for (int i = b0; i <= b1; i++) if (Ac->en.a[i] < emin) emin = Ac->en.a[i];
    if (emin > ec) { remove_block(Ac, b); b--; }
  }
  for (int b = 0; b < nblocks(Ac); b++) {
    int b0 = blim(Ac, b, 0), b1 = blim(Ac, b, 1);
    double emax = Ac->en.a[b0];
    for (int i = b0; i <= b1; i++) if (Ac->en.a[i] > emax) emax = Ac->en.a[i];
    if (emax > ec) {
      int ip = b0;
      while (ip <= b1 && Ac->en.a[ip] <= ec) ip++;
      remove_states_from_block(Ac, b, ip, b1);
    }
  }
}

/* f_ch in the eigenbasis: pairs (ibl < jbl) with Q_j = Q_i+1, S_j = S_i +- 1/2,
 * channel loop inside the pair loop.  A has BC data (bbe_bc/bbe_ev) when it is a
 * cut array; the matching pre-cut block of Ab is found by quantum numbers. */
static double qsz_fn_el(int ch, int ti, int tj);
static void twoch_update_f(const arr_t *S, const arr_t *A, const arr_t *Ab, mat_t *f /* [2] */, int qsz) {
  for (int c = 0; c < 2; c++) { mat_clear(&f[c]); mat_sync(&f[c], A); f[c].b.nshell = A->nshell; }
  int icount[2] = {0, 0};
  for (int ib = 0; ib < nblocks(A); ib++) {
    double Qi = getqn(A, ib, 0), Si = getqn(A, ib, 1);
    int ni = bsize(A, ib);
    double q[2] = {Qi, Si};
    int ibc = blk_from_qn(Ab, q);
    int nibc = bsize(Ab, ibc);
    for (int jb = ib + 1; jb < nblocks(A); jb++) {
      double Qj = getqn(A, jb, 0), Sj = getqn(A, jb, 1);
      int nj = bsize(A, jb);
      double qq[2] = {Qj, Sj};
      int jbc = blk_from_qn(Ab, qq);
      int njbc = bsize(Ab, jbc);
      if (!(dEqual(Qi + 1.0, Qj) && (dEqual(Si + 0.5, Sj) || dEqual(Si - 0.5, Sj)))) continue;
      for (int ch = 1; ch <= 2; ch++) {
        mat_t *M = &f[ch - 1];
        ipush(&M->map, ib); ipush(&M->map, jb);
        ipush(&M->mbe, icount[ch - 1]);
        icount[ch - 1] += ni * nj;
        ipush(&M->mbe, icount[ch - 1] - 1);
        double *F = (double *)calloc((size_t)nibc * (size_t)njbc, sizeof(double));
        for (int a = 0; a < nibc; a++) {
          int ist = blim(Ab, ibc, 0) + a;
          for (int b = 0; b < njbc; b++) {
            int jst = blim(Ab, jbc, 0) + b;
            if (Ab->stcf.a[ist] != Ab->stcf.a[jst]) continue;
            F[(size_t)a * njbc + b] = qsz ? qsz_fn_el(ch, Ab->itype.a[ist], Ab->itype.a[jst])
                                          : twoch_fn_el(S, ch, Si, Sj, Ab->itype.a[ist], Ab->itype.a[jst]);
          }
        }
        dres(&M->mat, M->mat.n + ni * nj);
        orc_rotate(nibc, njbc, ni, nj, A->ev.a + blim_ev(A, ib), F, A->ev.a + blim_ev(A, jb), M->mat.a + M->mat.n);
        M->mat.n += ni * nj;
        free(F);
      }
    }
  }
}

int orc_run_twochqs(const orc_twochqs_params *p, orc_step_cb cb, void *user) {
  const double L = p->Lambda, bbar = p->betabar > 0 ? p->betabar : 0.727;
  const double HLF = 0.5 * (1.0 + 1.0 / L);
  arr_t S = {0}, Ae = {0}, Ac = {0}, Ab = {0}, Ah = {0}, Ahc = {0};
  mat_t f[2], d = {0};
  memset(f, 0, sizeof(f));
  twochqs_site(&S);
  /* src/TwoChQS/TwoChQS.cpp:208-264 */
  double Ut = 0.5 * p->U / HLF, edt = p->ed / HLF;
  double G1t = p->Gamma1 * (2.0 / M_PI) / (HLF * HLF), G2t = p->Gamma2 * (2.0 / M_PI) / (HLF * HLF);
  double Un = Ut / sqrt(L), edn = edt / sqrt(L), chi0[2] = {sqrt(G1t / L), sqrt(G2t / L)};
  /* TwoChQS_SetH0Anderson  src/TwoChQS/TwoChQS_SetH0Anderson.cpp:15-345 */
  Ah.nq = 2; Ah.nshell = -1;
  const double qh[6] = {-1, 0, 0, 0.5, 1, 0};
  for (int i = 0; i < 6; i++) dpush(&Ah.qn, qh[i]);
  for (int i = 0; i < 3; i++) { ipush(&Ah.bbe, i); ipush(&Ah.bbe, i); dpush(&Ah.ev, 1.0); }
  dpush(&Ah.en, Un); dpush(&Ah.en, edn + Un); dpush(&Ah.en, 2.0 * edn + 3.0 * Un);
  mat_sync(&d, &Ah);
  dpush(&d.mat, 1.0); ipush(&d.map, 0); ipush(&d.map, 1); ipush(&d.mbe, 0); ipush(&d.mbe, 0);
  dpush(&d.mat, -sqrt(2.0)); ipush(&d.map, 1); ipush(&d.map, 2); ipush(&d.mbe, 1); ipush(&d.mbe, 1);
  cut_states_legacy(&Ah, &Ahc, 100);
  qs_build_basis(&Ahc, &Ab, &S, 1);
  Ae.nshell = Ab.nshell; Ae.nq = 2;
  arr_clear(&Ae);
  dcopy_v(&Ae.qn, &Ab.qn); icopy_v(&Ae.bbe, &Ab.bbe);
  for (int b = 0; b < nblocks(&Ab); b++) {
    int n = bsize(&Ab, b), b0 = blim(&Ab, b, 0);
    double Si = getqn(&Ab, b, 1);
    double *H = (double *)calloc((size_t)n * n, sizeof(double));
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) {
        int ist = b0 + i, jst = b0 + j;
        if (i == j) { H[(size_t)i * n + j] = Ab.en.a[ist]; continue; }
        double old = get_matel(&d, Ab.stcf.a[ist], Ab.stcf.a[jst]);
        int tp = Ab.itype.a[ist], tt = Ab.itype.a[jst];
        if (dEqual(old, 0.0)) { old = get_matel(&d, Ab.stcf.a[jst], Ab.stcf.a[ist]); tp = Ab.itype.a[jst]; tt = Ab.itype.a[ist]; }
        double v = 0.0;
        for (int ch = 1; ch <= 2; ch++) v += chi0[ch - 1] * old * twoch_hn_coef(&S, ch, Si, tp, tt, 1);
        H[(size_t)i * n + j] = v;
      }
    dres(&Ae.en, Ae.en.n + n); dres(&Ae.ev, Ae.ev.n + n * n);
    orc_eigh_full(n, H, Ae.en.a + Ae.en.n, Ae.ev.a + Ae.ev.n);
    Ae.en.n += n; Ae.ev.n += n * n;
    free(H);
  }
  double emin = set_e0zero(&Ae);
  twoch_update_f(&S, &Ae, &Ab, f, 0);   /* TwoChQS_UpdateQm1fQ on the uncut N=0 basis */
  double DN = HLF * pow(L, -(0 - 1) / 2.0);
  {
    double thT[4] = {DN / bbar, 0, 0, 0}, sus[4] = {calc_suscep(bbar, &Ae), 0, 0, 0};
    double ent[4] = {calc_entropy(bbar, &Ae), 0, 0, 0};
    arr_t tmpc = {0};
    cut_states_legacy(&Ae, &tmpc, p->ncut);
    emit(cb, user, &Ae, &tmpc, emin, HUGE_VAL, Ae.en.n, NULL, 0, thT, sus, ent, DN / bbar, 0, NULL);
    arr_free(&tmpc);
  }
  dvec hrec = {0};
  for (int N = 1; N <= p->nsitesmax; N++) {
    DN = HLF * pow(L, -(N - 1) / 2.0);
    double d0 = 1.0 - pow(L, -N), d1 = sqrt(1.0 - pow(L, -(2 * N - 1))), d2 = sqrt(1.0 - pow(L, -(2 * N + 1)));
    double chiN = d0 / (d1 * d2);
    cut_states_legacy(&Ae, &Ac, p->ncut);
    if (N > 1) twoch_update_f(&S, &Ac, &Ab, f, 0);
    qs_build_basis(&Ac, &Ab, &S, 0);
    /* TwoChQS_DiagHN */
    Ae.nshell = Ab.nshell; Ae.nq = 2;
    arr_clear(&Ae);
    dcopy_v(&Ae.qn, &Ab.qn); icopy_v(&Ae.bbe, &Ab.bbe);
    hrec.n = 0;
    for (int b = 0; b < nblocks(&Ab); b++) {
      int n = bsize(&Ab, b), b0 = blim(&Ab, b, 0);
      double Si = getqn(&Ab, b, 1);
      double *ap = (double *)malloc(sizeof(double) * (size_t)n * (n + 1) / 2);
      int r = 0;
      for (int i = 0; i < n; i++)
        for (int j = i; j < n; j++) {
          int ist = b0 + i, jst = b0 + j;
          double v;
          if (i == j) v = sqrt(L) * Ab.en.a[ist];
          else {
            v = 0.0;
            for (int ch = 1; ch <= 2; ch++) {
              double old = get_matel(&f[ch - 1], Ab.stcf.a[ist], Ab.stcf.a[jst]);
              int tp = Ab.itype.a[ist], tt = Ab.itype.a[jst];
              if (dEqual(fabs(old), 0.0)) {
                old = get_matel(&f[ch - 1], Ab.stcf.a[jst], Ab.stcf.a[ist]);
                tp = Ab.itype.a[jst]; tt = Ab.itype.a[ist];
              }
              v += chiN * old * twoch_hn_coef(&S, ch, Si, tp, tt, 0);
            }
          }
          ap[r++] = v;
        }
      if (p->record_blocks) for (int k = 0; k < r; k++) dpush(&hrec, ap[k]);
      dres(&Ae.en, Ae.en.n + n); dres(&Ae.ev, Ae.ev.n + n * n);
      orc_eigh_packed(n, ap, Ae.en.a + Ae.en.n, Ae.ev.a + Ae.ev.n);
      Ae.en.n += n; Ae.ev.n += n * n;
      free(ap);
    }
    emin = set_e0zero(&Ae);
    double thT[4] = {DN / bbar, 0, 0, 0}, sus[4] = {calc_suscep(bbar, &Ae), 0, 0, 0};
    double ent[4] = {calc_entropy(bbar, &Ae), 0, 0, 0};
    /* record the cut that the next iteration will make */
    arr_t nc = {0};
    double ec = HUGE_VAL;
    if (p->ncut < nstates(&Ae)) ec = orc_ecut(Ae.en.n, Ae.en.a, p->ncut);
    cut_states_legacy(&Ae, &nc, p->ncut);
    int nk = 0;
    Ae.kept.n = 0;
    for (int i = 0; i < Ae.en.n; i++) { int k = Ae.en.a[i] <= ec; ipush(&Ae.kept, k); nk += k; }
    emit(cb, user, &Ae, &nc, emin, ec, nk, &hrec, p->record_blocks, thT, sus, ent, DN / bbar, 0, NULL);
    arr_free(&nc);
  }
  free(hrec.a);
  mat_free(&f[0]); mat_free(&f[1]); mat_free(&d);
  arr_free(&S); arr_free(&Ae); arr_free(&Ac); arr_free(&Ab); arr_free(&Ah); arr_free(&Ahc);
  return 0;
}

/* ================================================================== */
/* TwoChQSz Anderson: the standalone driver src/TwoChQSz/TwoChQSz.cpp  */
/* :103-425 (the part after its self-test exit): (Q, Sz) blocks, no    */
/* SU(2); QSz_BuildBasis, TwoChQSz_DiagHN, TwoChQSz_SetH0Anderson,      */
/* TwoChQSz_UpdateQm1fQ / _UpdateMatrixAfterCutting, CalcSuscep with    */
/* totalS false.                                                        */
/* ================================================================== */

/* TwoChQSz_SetSingleSite  src/TwoChQSz/TwoChQSz_SetSingleSite.cpp:5-85: 16 states,
 * each its own (Q, Sz) block, iType = state index */
static void twochqsz_site(arr_t *S) {
  arr_clear(S);
  S->nq = 2; S->nshell = 0;
  const double q[32] = {-2, 0, -1, 0.5, -1, -0.5, 0, 0, -1, 0.5, 0, 1, 0, 0, 1, 0.5,
                        -1, -0.5, 0, 0, 0, -1, 1, -0.5, 0, 0, 1, 0.5, 1, -0.5, 2, 0};
  for (int i = 0; i < 32; i++) dpush(&S->qn, q[i]);
  for (int i = 0; i < 16; i++) { ipush(&S->bbe, i); ipush(&S->bbe, i); ipush(&S->itype, i); }
}
/* fd_table  src/TwoChQSz/TwoChQSz_OpTables.cpp:3-78: <type_i| f^dagger_{ch,sigma} |type_j>;
 * site state bits: 1 = ch2 up, 2 = ch2 down, 4 = ch1 up, 8 = ch1 down */
static double qsz_fd(int ch, int sigma, int ti, int tj) {
  if (ti < 0 || ti > 15 || tj < 0 || tj > 15) return 0.0;
  int bit = (ch == 1) ? (sigma == 1 ? 4 : (sigma == -1 ? 8 : 0)) : (ch == 2 ? (sigma == 1 ? 1 : (sigma == -1 ? 2 : 0)) : 0);
  if (!bit || (tj & bit) || ti != (tj | bit)) return 0.0;
  /* the signs of the table: ch1 up +1; ch1 dn -1 when ch1 up is occupied; ch2 up
   * -1 when exactly one of the ch1 modes is occupied; ch2 dn: that sign times
   * -1 when ch2 up is occupied */
  if (bit == 4) return 1.0;
  if (bit == 8) return (tj & 4) ? -1.0 : 1.0;
  double s = (((tj >> 2) & 1) ^ ((tj >> 3) & 1)) ? -1.0 : 1.0;
  if (bit == 1) return s;
  return (tj & 1) ? -s : s;
}
/* fnbasis of TwoChQSz_UpdateQm1fQ / _UpdateMatrixAfterCutting
 * (src/TwoChQSz/TwoChQSz_UpdateMatrices.cpp:110-111, 288-289), same parent */
static double qsz_fn_el(int ch, int ti, int tj) { return qsz_fd(ch, -1, tj, ti) + qsz_fd(ch, 1, tj, ti); }
/* H_N coupling of TwoChQSz_DiagHN (src/TwoChQSz/TwoChQSz_DiagHN.cpp:74-97, h0 = 0)
 * and TwoChQSz_SetH0Anderson (src/TwoChQSz/TwoChQSz_SetH0Anderson.cpp:180-215,
 * h0 = 1): sum over sigma of fd(ch, sigma, typep, type) times the fermion sign
 * of the site state `type` (-1 for Sz = +-1/2; at N = 0 also Q = +-1) */
static double qsz_hn_coef(const arr_t *S, int ch, int typep, int type, int h0) {
  double Qt = getqn(S, type, 0), Szt = getqn(S, type, 1);
  double fs = 1.0;
  if (h0) { if ((dEqual(Qt, 1.0) || dEqual(Qt, -1.0)) && (dEqual(Szt, 0.5) || dEqual(Szt, -0.5))) fs = -1.0; }
  else if (dEqual(Szt, 0.5) || dEqual(Szt, -0.5)) fs = -1.0;
  double v = 0.0;
  for (int sg = -1; sg <= 1; sg += 2) v += qsz_fd(ch, sg, typep, type) * fs;
  return v;
}
/* QSz_BuildBasis  src/QSz_BuildBasis.cpp:9-117: (Q, Sz) of old block + site
 * block, FilterQNumbers (first appearance), then per new block: old block,
 * site block, old state */
static void qsz_build_basis(const arr_t *Ac, arr_t *Ab, const arr_t *S, int befcut) {
  Ab->nshell = Ac->nshell + 1;
  Ab->nq = 2;
  arr_clear(Ab);
  for (int b = 0; b < nblocks(Ac); b++)
    for (int s = 0; s < nblocks(S); s++) {
      dpush(&Ab->qn, getqn(Ac, b, 0) + getqn(S, s, 0));
      dpush(&Ab->qn, getqn(Ac, b, 1) + getqn(S, s, 1));
    }
  filter_qnumbers(Ab);
  int istnew = 0;
  for (int bn = 0; bn < nblocks(Ab); bn++) {
    double Qn = getqn(Ab, bn, 0), Sn = getqn(Ab, bn, 1);
    ipush(&Ab->bbe, istnew);
    for (int bo = 0; bo < nblocks(Ac); bo++) {
      double Qo = getqn(Ac, bo, 0), So = getqn(Ac, bo, 1);
      for (int s = 0; s < nblocks(S); s++) {
        if (!(dEqual(Qn, Qo + getqn(S, s, 0)) && dEqual(Sn, So + getqn(S, s, 1)))) continue;
        for (int io = blim(Ac, bo, 0); io <= blim(Ac, bo, 1); io++) {
          ipush(&Ab->itype, S->itype.a[s]);
          dpush(&Ab->en, Ac->en.a[io]);
          ipush(&Ab->stcf, befcut == 1 ? Ac->stcf.a[io] : io);
          istnew++;
        }
      }
    }
    ipush(&Ab->bbe, istnew - 1);
  }
}
/* CalcSuscep / CalcEntropy with totalS false (src/CalcThermo.cpp:16-121):
 * <Sz^2> and S with degeneracy 1 per state */
static double calc_suscep_sz(double bb, const arr_t *A) {
  double s0 = 0, s1 = 0;
  for (int b = 0; b < nblocks(A); b++) {
    double Sz = getqn(A, b, 1);
    for (int i = blim(A, b, 0); i <= blim(A, b, 1); i++) {
      double x = exp(-bb * A->en.a[i]);
      s0 += x; s1 += Sz * Sz * x;
    }
  }
  return s1 / s0;
}
static double calc_entropy_sz(double bb, const arr_t *A) {
  double Z = 0, s = 0;
  for (int i = 0; i < A->en.n; i++) { double E = A->en.a[i], x = exp(-bb * E); s += E * x; Z += x; }
  return bb * (s / Z) + log(Z);
}

/* d_up on every state of a shell (for the DM-NRG post-processing of TwoChQSz,
 * which the reference does not have): for each block pair (i, j) of Aeig with
 * Q_j = Q_i + 1 and Sz_j = Sz_i + 1/2, Z_i F Z_j^T over all eigenvector rows,
 * F(ist, jst) = delta(type_ist, type_jst) (-1)^{n(type)} d_old(stcf_ist, stcf_jst):
 * d acts on the old part, past the new site's n(type) fermions (the ordering of
 * the coupling signs of qsz_hn_coef). */
static void qsz_dup_all(const arr_t *Ae, const arr_t *Ab, const mat_t *dold, mat_t *D) {
  mat_clear(D);
  mat_sync(D, Ae);
  int icount = 0;
  for (int ib = 0; ib < nblocks(Ae); ib++)
    for (int jb = 0; jb < nblocks(Ae); jb++) {
      if (!(dEqual(getqn(Ae, jb, 0), getqn(Ae, ib, 0) + 1.0) && dEqual(getqn(Ae, jb, 1), getqn(Ae, ib, 1) + 0.5)))
        continue;
      int ni = bsize(Ae, ib), nj = bsize(Ae, jb);
      ipush(&D->map, ib); ipush(&D->map, jb);
      ipush(&D->mbe, icount);
      icount += ni * nj;
      ipush(&D->mbe, icount - 1);
      double *F = (double *)calloc((size_t)ni * nj + 1, sizeof(double));
      for (int a = 0; a < ni; a++) {
        int ist = blim(Ab, ib, 0) + a, t = Ab->itype.a[ist];
        for (int b = 0; b < nj; b++) {
          int jst = blim(Ab, jb, 0) + b;
          if (Ab->itype.a[jst] != t) continue;
          double sg = (__builtin_popcount((unsigned)t) & 1) ? -1.0 : 1.0;
          F[(size_t)a * nj + b] = sg * get_matel(dold, Ab->stcf.a[ist], Ab->stcf.a[jst]);
        }
      }
      dres(&D->mat, D->mat.n + ni * nj);
      orc_rotate(ni, nj, ni, nj, Ae->ev.a + blim_ev(Ae, ib), F, Ae->ev.a + blim_ev(Ae, jb), D->mat.a + D->mat.n);
      D->mat.n += ni * nj;
      free(F);
    }
}
/* the kept part of an all-states operator: the kept prefixes of the surviving blocks */
static void op_kept_part(const mat_t *Dall, const arr_t *Ae, const arr_t *Ac, mat_t *Dk) {
  mat_clear(Dk);
  mat_sync(Dk, Ac);
  int icount = 0;
  for (int ib = 0; ib < nblocks(Ac); ib++)
    for (int jb = 0; jb < nblocks(Ac); jb++) {
      int bi = blk_from_qn(Ae, Ac->qn.a + ib * Ac->nq), bj = blk_from_qn(Ae, Ac->qn.a + jb * Ac->nq);
      int im = find_mat_block(Dall, bi, bj);
      if (im < 0) continue;
      int ki = bsize(Ac, ib), kj = bsize(Ac, jb), nj = bsize(Ae, bj);
      ipush(&Dk->map, ib); ipush(&Dk->map, jb);
      ipush(&Dk->mbe, icount);
      icount += ki * kj;
      ipush(&Dk->mbe, icount - 1);
      for (int a = 0; a < ki; a++)
        for (int b = 0; b < kj; b++) dpush(&Dk->mat, Dall->mat.a[Dall->mbe.a[2 * im] + a * nj + b]);
    }
}
static void dm_store_qsz(struct orc_dm *D, const arr_t *Ae, const arr_t *Ab, const arr_t *Acprev, const mat_t *dall);
static int run_twochqsz(const orc_twochqs_params *p, orc_step_cb cb, void *user, struct orc_dm *dm);
int orc_run_twochqsz(const orc_twochqs_params *p, orc_step_cb cb, void *user) { return run_twochqsz(p, cb, user, NULL); }
/* dm != NULL: every shell (N = 0 .. nsitesmax) is stored with all its states,
 * their Kept flags and d_up on them, Abasis with StCameFrom as indices of the
 * previous shell's states before the cut */
static int run_twochqsz(const orc_twochqs_params *p, orc_step_cb cb, void *user, struct orc_dm *dm) {
  const double L = p->Lambda, bbar = p->betabar > 0 ? p->betabar : 0.727;
  const double HLF = 0.5 * (1.0 + 1.0 / L);
  arr_t S = {0}, Ae = {0}, Ac = {0}, Ab = {0}, Ah = {0}, Ahc = {0};
  mat_t f[2], d = {0};
  memset(f, 0, sizeof(f));
  twochqsz_site(&S);
  /* src/TwoChQSz/TwoChQSz.cpp:156-185 */
  double Ut = 0.5 * p->U / HLF, edt = p->ed / HLF;
  double G1t = p->Gamma1 * (2.0 / M_PI) / (HLF * HLF), G2t = p->Gamma2 * (2.0 / M_PI) / (HLF * HLF);
  double Un = Ut / sqrt(L), edn = edt / sqrt(L), chi0[2] = {sqrt(G1t / L), sqrt(G2t / L)};
  /* TwoChQSz_SetH0Anderson  src/TwoChQSz/TwoChQSz_SetH0Anderson.cpp:15-251: four
   * impurity blocks (Q, Sz) and the impurity operator of its :90-135 */
  Ah.nq = 2; Ah.nshell = -1;
  const double qh[8] = {-1, 0, 0, 0.5, 0, -0.5, 1, 0};
  for (int i = 0; i < 8; i++) dpush(&Ah.qn, qh[i]);
  for (int i = 0; i < 4; i++) { ipush(&Ah.bbe, i); ipush(&Ah.bbe, i); dpush(&Ah.ev, 1.0); }
  dpush(&Ah.en, Un); dpush(&Ah.en, edn + Un); dpush(&Ah.en, edn + Un); dpush(&Ah.en, 2.0 * edn + 3.0 * Un);
  mat_sync(&d, &Ah);
  const int dmap[8] = {0, 1, 0, 2, 1, 3, 2, 3};
  const double dval[4] = {1.0, 1.0, -1.0, 1.0};
  for (int k = 0; k < 4; k++) {
    dpush(&d.mat, dval[k]); ipush(&d.map, dmap[2 * k]); ipush(&d.map, dmap[2 * k + 1]);
    ipush(&d.mbe, k); ipush(&d.mbe, k);
  }
  cut_states_legacy(&Ah, &Ahc, 100);
  qsz_build_basis(&Ahc, &Ab, &S, 1);
  Ae.nshell = Ab.nshell; Ae.nq = 2;
  arr_clear(&Ae);
  dcopy_v(&Ae.qn, &Ab.qn); icopy_v(&Ae.bbe, &Ab.bbe);
  for (int b = 0; b < nblocks(&Ab); b++) {
    int n = bsize(&Ab, b), b0 = blim(&Ab, b, 0);
    double *H = (double *)calloc((size_t)n * n, sizeof(double));
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) {
        int ist = b0 + i, jst = b0 + j;
        if (i == j) { H[(size_t)i * n + j] = Ab.en.a[ist]; continue; }
        double old = get_matel(&d, Ab.stcf.a[ist], Ab.stcf.a[jst]);
        int tp = Ab.itype.a[ist], tt = Ab.itype.a[jst];
        if (dEqual(old, 0.0)) { old = get_matel(&d, Ab.stcf.a[jst], Ab.stcf.a[ist]); tp = Ab.itype.a[jst]; tt = Ab.itype.a[ist]; }
        double v = 0.0;
        for (int ch = 1; ch <= 2; ch++) v += chi0[ch - 1] * old * qsz_hn_coef(&S, ch, tp, tt, 1);
        H[(size_t)i * n + j] = v;
      }
    dres(&Ae.en, Ae.en.n + n); dres(&Ae.ev, Ae.ev.n + n * n);
    orc_eigh_full(n, H, Ae.en.a + Ae.en.n, Ae.ev.a + Ae.ev.n);
    Ae.en.n += n; Ae.ev.n += n * n;
    free(H);
  }
  double emin = set_e0zero(&Ae);
  twoch_update_f(&S, &Ae, &Ab, f, 1);   /* TwoChQSz_UpdateQm1fQ on the uncut N=0 basis */
  mat_t dup = {0}, dall = {0};          /* d_up: impurity table (up: (0,1) and (2,3)), then per shell */
  if (dm) {
    mat_sync(&dup, &Ah);
    const int um[4] = {0, 1, 2, 3};
    for (int k = 0; k < 2; k++) {
      dpush(&dup.mat, 1.0); ipush(&dup.map, um[2 * k]); ipush(&dup.map, um[2 * k + 1]);
      ipush(&dup.mbe, k); ipush(&dup.mbe, k);
    }
    qsz_dup_all(&Ae, &Ab, &dup, &dall);
    arr_t kc = {0};
    cut_states_legacy(&Ae, &kc, p->ncut);
    Ae.kept.n = 0;
    double ec0 = (p->ncut < nstates(&Ae)) ? orc_ecut(Ae.en.n, Ae.en.a, p->ncut) : HUGE_VAL;
    for (int i = 0; i < Ae.en.n; i++) ipush(&Ae.kept, Ae.en.a[i] <= ec0);
    dm_store_qsz(dm, &Ae, &Ab, &Ahc, &dall);
    op_kept_part(&dall, &Ae, &kc, &dup);
    arr_free(&kc);
  }
  double DN = HLF * pow(L, -(0 - 1) / 2.0);
  {
    double thT[4] = {DN / bbar, 0, 0, 0}, sus[4] = {calc_suscep_sz(bbar, &Ae), 0, 0, 0};
    double ent[4] = {calc_entropy_sz(bbar, &Ae), 0, 0, 0};
    arr_t tmpc = {0};
    cut_states_legacy(&Ae, &tmpc, p->ncut);
    emit(cb, user, &Ae, &tmpc, emin, HUGE_VAL, Ae.en.n, NULL, 0, thT, sus, ent, DN / bbar, 0, NULL);
    arr_free(&tmpc);
  }
  dvec hrec = {0};
  for (int N = 1; N <= p->nsitesmax; N++) {
    DN = HLF * pow(L, -(N - 1) / 2.0);
    double d0 = 1.0 - pow(L, -N), d1 = sqrt(1.0 - pow(L, -(2 * N - 1))), d2 = sqrt(1.0 - pow(L, -(2 * N + 1)));
    double chiN = d0 / (d1 * d2);
    cut_states_legacy(&Ae, &Ac, p->ncut);
    if (N > 1) twoch_update_f(&S, &Ac, &Ab, f, 1);
    if (dm && N > 1) op_kept_part(&dall, &Ae, &Ac, &dup);
    qsz_build_basis(&Ac, &Ab, &S, 0);
    /* TwoChQSz_DiagHN: the rules give a symmetric H; the upper triangle is used */
    Ae.nshell = Ab.nshell; Ae.nq = 2;
    arr_clear(&Ae);
    dcopy_v(&Ae.qn, &Ab.qn); icopy_v(&Ae.bbe, &Ab.bbe);
    hrec.n = 0;
    for (int b = 0; b < nblocks(&Ab); b++) {
      int n = bsize(&Ab, b), b0 = blim(&Ab, b, 0);
      double *ap = (double *)malloc(sizeof(double) * ((size_t)n * (n + 1) / 2 + 1));
      int r = 0;
      for (int i = 0; i < n; i++)
        for (int j = i; j < n; j++) {
          int ist = b0 + i, jst = b0 + j;
          double v;
          if (i == j) v = sqrt(L) * Ab.en.a[ist];
          else {
            v = 0.0;
            for (int ch = 1; ch <= 2; ch++) {
              double old = get_matel(&f[ch - 1], Ab.stcf.a[ist], Ab.stcf.a[jst]);
              int tp = Ab.itype.a[ist], tt = Ab.itype.a[jst];
              if (dEqual(fabs(old), 0.0)) {
                old = get_matel(&f[ch - 1], Ab.stcf.a[jst], Ab.stcf.a[ist]);
                tp = Ab.itype.a[jst]; tt = Ab.itype.a[ist];
              }
              v += chiN * old * qsz_hn_coef(&S, ch, tp, tt, 0);
            }
          }
          ap[r++] = v;
        }
      if (p->record_blocks) for (int k = 0; k < r; k++) dpush(&hrec, ap[k]);
      dres(&Ae.en, Ae.en.n + n); dres(&Ae.ev, Ae.ev.n + n * n);
      orc_eigh_packed(n, ap, Ae.en.a + Ae.en.n, Ae.ev.a + Ae.ev.n);
      Ae.en.n += n; Ae.ev.n += n * n;
      free(ap);
    }
    emin = set_e0zero(&Ae);
    double thT[4] = {DN / bbar, 0, 0, 0}, sus[4] = {calc_suscep_sz(bbar, &Ae), 0, 0, 0};
    double ent[4] = {calc_entropy_sz(bbar, &Ae), 0, 0, 0};
    arr_t nc = {0};
    double ec = HUGE_VAL;
    if (p->ncut < nstates(&Ae)) ec = orc_ecut(Ae.en.n, Ae.en.a, p->ncut);
    cut_states_legacy(&Ae, &nc, p->ncut);
    int nk = 0;
    Ae.kept.n = 0;
    for (int i = 0; i < Ae.en.n; i++) { int k = Ae.en.a[i] <= ec; ipush(&Ae.kept, k); nk += k; }
    if (dm) {
      qsz_dup_all(&Ae, &Ab, &dup, &dall);
      dm_store_qsz(dm, &Ae, &Ab, &Ac, &dall);
    }
    emit(cb, user, &Ae, &nc, emin, ec, nk, &hrec, p->record_blocks, thT, sus, ent, DN / bbar, 0, NULL);
    arr_free(&nc);
  }
  free(hrec.a);
  mat_free(&dup); mat_free(&dall);
  mat_free(&f[0]); mat_free(&f[1]); mat_free(&d);
  arr_free(&S); arr_free(&Ae); arr_free(&Ac); arr_free(&Ab); arr_free(&Ah); arr_free(&Ahc);
  return 0;
}

/* ================================================================== */
/* DM-NRG: density-matrix sweep and spectral sums (src/DM_NRG,         */
/* src/SpecFuncClass.cpp, src/SpecFunc_allSTL.cpp, src/BroadDelta.cpp)  */
/* for the OneChQS Anderson chain run with calcdens 1 (ops f_N, c_d),   */
/* UpdateBefCut 0 or 1; DM-NRG (UseCFS 0), CFS (-C) or FDM (-F).        */
/* ================================================================== */
struct orc_dm {
  orc_dmnrg_params p;
  int nsh;                 /* shells stored: 0 .. nsh-1 (Nsites0 = 0) */
  arr_t *acut, *abasis;    /* Acut_N, Abasis_N as SaveArrays writes them */
  mat_t *op;               /* Mat0_N = c_d (cdot1) in the Acut_N basis */
  mat_t *rho;              /* rho_N, packed upper diagonal blocks over Acut_N */
  double *trace;           /* Trace (2S+1) rho_N printed by CalcRhoN / SetRhoNmax */
  int nshell_max;          /* DM_NRG NshellMax */
  int spec_nshell_max;     /* CSpecFunction::NshellMax */
  double temp, betabar_spec;
  int method;              /* spec1.UseCFS: 0 DM-NRG, 1 CFS, 2 FDM */
  double ztot;             /* FDM: Z_tot (src/DM_NRG/DM_NRG.cpp:203-268) */
  double *wn;              /* FDM: w_N per shell */
  int fix;                 /* FDM: ORC_FDM_FIX_* bits in effect (0 unless FDM) */
  double bb_last;          /* FDM with ORC_FDM_FIX_LAST: 1/TempBar of the last shell */
  arr_t site;
  int nomega;
  double *omega, *rhow;
  /* derivative pass (method 0): per shell the kept-space tangents of the chain
   * (th per kept block, c_d, kept rows dZ) and of rho_N, per tangent t */
  int ntan;
  dvec (*th)[ORC_MAX_TAN];
  ivec *th_off;
  dvec (*dop)[ORC_MAX_TAN];
  dvec (*dz)[ORC_MAX_TAN];
  dvec (*drho)[ORC_MAX_TAN];
};

static void arr_copy(arr_t *d, const arr_t *s) {
  d->nshell = s->nshell; d->nq = s->nq; d->sz = s->sz;
  dcopy_v(&d->qn, &s->qn); icopy_v(&d->bbe, &s->bbe); dcopy_v(&d->en, &s->en); dcopy_v(&d->ev, &s->ev);
  icopy_v(&d->kept, &s->kept); icopy_v(&d->itype, &s->itype); icopy_v(&d->stcf, &s->stcf);
  icopy_v(&d->bbe_bc, &s->bbe_bc); icopy_v(&d->bbe_ev, &s->bbe_ev);
}

/* CalcDN  src/CalcDN.cpp:4-15 (z_twist = 1) */
static double calc_dn(double L, int N) {
  double HLF = (L > 0.0) ? 0.5 * (1.0 + 1.0 / L) : 1.0;
  return HLF * pow(L, -(N - 1) / 2.0);
}
/* CNRGarray::PartitionFunc  src/NRGarrayClass.cpp:997-1030 ((2S+1) per state, S = QN 1) */
/* degeneracy of a state of block b: 2S+1 (SU(2) blocks), 1 (Sz blocks) */
static double blk_deg(const arr_t *A, int b) { return A->sz ? 1.0 : 2.0 * getqn(A, b, 1) + 1.0; }
static double partition_func(const arr_t *A, double bb) {
  double s = 0.0;
  for (int b = 0; b < nblocks(A); b++)
    for (int i = blim(A, b, 0); i <= blim(A, b, 1); i++) s += blk_deg(A, b) * exp(-bb * A->en.a[i]);
  return s;
}
/* CNRGarray::PartitionFuncDisc  src/NRGarrayClass.cpp:1032-1058: discarded states only */
static double partition_func_disc(const arr_t *A, double bb) {
  if (A->kept.n == 0) return -1.0;
  double s = 0.0;
  for (int b = 0; b < nblocks(A); b++)
    for (int i = blim(A, b, 0); i <= blim(A, b, 1); i++)
      if (!A->kept.a[i]) s += blk_deg(A, b) * exp(-bb * A->en.a[i]);
  return s;
}
/* CNRGarray::CheckKept(ist, kp)  src/NRGarrayClass.cpp:493-503 */
static int check_kept(const arr_t *A, int ist, int kp) {
  if (A->kept.n <= ist) return 0;
  return (A->kept.a[ist] != 0) == (kp != 0);
}
/* CNRGarray::GetBlockSize(ibl, kp)  src/NRGarrayClass.cpp:357-371 */
static int bsize_kp(const arr_t *A, int b, int kp) {
  if (A->kept.n <= blim(A, b, 1)) return 0;
  int n = 0;
  for (int i = blim(A, b, 0); i <= blim(A, b, 1); i++) n += check_kept(A, i, kp);
  return n;
}
/* BroadDelta  src/BroadDelta.cpp:12-29 */
static double broad_delta(double w, double Ep, double b) {
  if (Ep == 0.0) Ep = 1.0e-100;
  if (w * Ep < 0.0) return 0.0;
  double a0 = exp(-b * b / 4.0) / (b * fabs(Ep) * sqrt(M_PI));
  double a1 = log(fabs(w / Ep)) / b;
  return a0 * exp(-a1 * a1);
}
/* LorentzDelta  src/BroadDelta.cpp:63-73 */
static double lorentz_delta(double w, double Ep, double b) {
  return (1.0 / (2.0 * M_PI)) * (b / ((w - Ep) * (w - Ep) + b * b));
}
/* LorentzDeltaAnders  src/BroadDelta.cpp:76-87 */
static double lorentz_delta_anders(double w, double Ep, double b) {
  return (1.0 / M_PI) * (b / ((w - Ep) * (w - Ep) + b * b));
}
/* GaussDelta  src/BroadDelta.cpp:96-107 */
static double gauss_delta(double w, double Ep, double b) {
  double x = (w - Ep) / b;
  return (1.0 / (b * sqrt(M_PI))) * exp(-x * x);
}
/* LogGaussDelta  src/BroadDelta.cpp:116-132 (Weichselbaum) */
static double loggauss_delta(double w, double Ep, double b) {
  if (Ep == 0.0) Ep = 1.0e-100;
  if (w * Ep < 0.0) return 0.0;
  double a0 = 1.0 / (b * fabs(Ep) * sqrt(M_PI));
  double a1 = log(fabs(w / Ep)) / b;
  return a0 * exp(-(a1 - b / 4.0) * (a1 - b / 4.0));
}

/* DM_NRG_SetRhoNmax  src/DM_NRG/DM_NRG_SetRhoNmax.cpp:9-126: packed upper,
 * block diagonal, diag exp(-betabar E)/ZN */
static double set_rho_nmax(double TempBar, double ZN, const arr_t *A, mat_t *R) {
  mat_clear(R);
  mat_sync(R, A);
  R->upper = 1;
  double bb = (fabs(TempBar) < 1e-20) ? 1.0e10 : 1.0 / TempBar;
  int i1 = 0;
  double tr = 0.0;
  for (int b = 0; b < nblocks(A); b++) {
    ipush(&R->map, b); ipush(&R->map, b); ipush(&R->mbe, i1);
    int i0 = blim(A, b, 0), i9 = blim(A, b, 1);
    for (int i = i0; i <= i9; i++) {
      double x = exp(-bb * A->en.a[i]) / ZN;
      dpush(&R->mat, x); i1++;
      tr += blk_deg(A, b) * x;
      for (int j = i + 1; j <= i9; j++) { dpush(&R->mat, 0.0); i1++; }
    }
    ipush(&R->mbe, i1 - 1);
  }
  return tr;
}

/* RotateMatrix  src/RotateMatrix.cpp:17-99 with RotateBlock2BLAS
 * src/NRGmatrixClass.cpp:1471-1512: M_rot = Z1^T (M Z2) per matrix block, Z = kept
 * eigenvector rows (k x n_BC); the result lives on BlockBegEndBC */
static void rotate_matrix(const mat_t *M, const arr_t *Ac, mat_t *R) {
  mat_clear(R);
  icopy_v(&R->map, &M->map);
  R->upper = M->upper;
  R->b.nq = Ac->nq; R->b.nshell = Ac->nshell;
  dcopy_v(&R->b.qn, &Ac->qn);
  if (Ac->bbe_bc.n) icopy_v(&R->b.bbe, &Ac->bbe_bc); else icopy_v(&R->b.bbe, &Ac->bbe);
  int i1 = 0;
  for (int im = 0; im < M->map.n / 2; im++) {
    int b1 = M->map.a[2 * im], b2 = M->map.a[2 * im + 1];
    ipush(&R->mbe, i1);
    int k1 = bsize(Ac, b1), k2 = bsize(Ac, b2), n1 = bsize_bc(Ac, b1), n2 = bsize_bc(Ac, b2);
    const double *Z1 = Ac->ev.a + Ac->bbe_ev.a[2 * b1], *Z2 = Ac->ev.a + Ac->bbe_ev.a[2 * b2];
    double *MZ = (double *)calloc((size_t)k1 * n2 + 1, sizeof(double));
    double *F = (double *)calloc((size_t)n1 * n2 + 1, sizeof(double));
    double *Mb = (double *)calloc((size_t)k1 * k2 + 1, sizeof(double));   /* MatBlock2BLAS */
    for (int i = 0; i < k1; i++) for (int j = 0; j < k2; j++) Mb[(size_t)i * k2 + j] = get_block_matel(M, b1, b2, i, j);
    for (int i = 0; i < k1; i++)        /* MZ = M (k1 x k2) * Z2 (k2 x n2) */
      for (int c = 0; c < n2; c++) {
        double s = 0.0;
        for (int j = 0; j < k2; j++) s += Mb[(size_t)i * k2 + j] * Z2[(size_t)j * n2 + c];
        MZ[(size_t)i * n2 + c] = s;
      }
    free(Mb);
    for (int r = 0; r < n1; r++)        /* F = Z1^T (n1 x k1) * MZ (k1 x n2) */
      for (int c = 0; c < n2; c++) {
        double s = 0.0;
        for (int i = 0; i < k1; i++) s += Z1[(size_t)i * n1 + r] * MZ[(size_t)i * n2 + c];
        F[(size_t)r * n2 + c] = s;
      }
    for (int r = 0; r < n1; r++)
      for (int c = (R->upper && b1 == b2) ? r : 0; c < n2; c++) { dpush(&R->mat, F[(size_t)r * n2 + c]); i1++; }
    ipush(&R->mbe, i1 - 1);
    free(MZ); free(F);
  }
}

/* DM_NRG_CalcRhoN_withSU2  src/DM_NRG/DM_NRG_CalcRhoN.cpp:252-449 with
 * DM_NRG_SetChildSt / DM_NRG_SetChildSameType src/DM_NRG/DM_NRG_SetChildSt.cpp:11-113:
 * rho_N(i,j) = sum over children (ci,cj) of the same site type and the same
 * N+1 block of CG(Sold,Sold; S~,Sz-Sold | Si,Sz)^2 summed over Sz, times
 * rho_rot(ci,cj).  FDM (:393-398): the diagonal of a discarded state is
 * exp(-E/TempBar)/ZNdisc instead (when ZNdisc != 0). */
/* GetMatEl on the rotated rho (block diagonal, packed upper): state -> block by table */
static double rot_el(const mat_t *R, const int *rblk, const int *rbeg, int ist, int jst) {
  int b1 = rblk[ist], b2 = rblk[jst];
  if (b1 < 0 || b2 < 0 || b1 != b2) return get_matel(R, ist, jst);
  int im = find_mat_block(R, b1, b1);
  if (im < 0) return 0.0;
  int n = bsize(&R->b, b1), i = ist - rbeg[b1], j = jst - rbeg[b1];
  int r = (i > j) ? ij2rUpTr(n, j, i) : ij2rUpTr(n, i, j);
  return R->mat.a[R->mbe.a[2 * im] + r];
}
static double gather_rho(const arr_t *AcN, const arr_t *AbNp1, const arr_t *S, const mat_t *rot, mat_t *RhoN,
                         int fdm, double TempBar, double ZNdisc);
static double calc_rho_n(const arr_t *AcN, const arr_t *AcNp1, const arr_t *AbNp1, const arr_t *S,
                         const mat_t *RhoNp1, mat_t *RhoN, int fdm, double TempBar, double ZNdisc) {
  mat_t rot = {0};
  rotate_matrix(RhoNp1, AcNp1, &rot);
  double tr = gather_rho(AcN, AbNp1, S, &rot, RhoN, fdm, TempBar, ZNdisc);
  mat_free(&rot);
  return tr;
}
/* the child trace-out of DM_NRG_CalcRhoN_withSU2 on an already rotated rho_{N+1} */
static double gather_rho(const arr_t *AcN, const arr_t *AbNp1, const arr_t *S, const mat_t *rotp, mat_t *RhoN,
                         int fdm, double TempBar, double ZNdisc) {
  const mat_t rot = *rotp;
  int nst1 = AbNp1->stcf.n;
  int *rblk = (int *)malloc(sizeof(int) * (size_t)(nst1 + 1));
  int *rbeg = (int *)malloc(sizeof(int) * (size_t)(nblocks(&rot.b) + 1));
  for (int i = 0; i < nst1; i++) rblk[i] = -1;
  for (int b = 0; b < nblocks(&rot.b); b++) {
    rbeg[b] = blim(&rot.b, b, 0);
    for (int i = blim(&rot.b, b, 0); i <= blim(&rot.b, b, 1) && i < nst1; i++) rblk[i] = b;
  }
  int ns = AcN->en.n;
  /* ChildStates: N+1 basis indices per Acut_N state, in basis order */
  int *ccount = (int *)calloc((size_t)ns + 1, sizeof(int)), *cbeg = (int *)calloc((size_t)ns + 2, sizeof(int));
  for (int c = 0; c < AbNp1->stcf.n; c++) ccount[AbNp1->stcf.a[c]]++;
  for (int i = 0; i < ns; i++) cbeg[i + 1] = cbeg[i] + ccount[i];
  int *child = (int *)malloc(sizeof(int) * (size_t)(cbeg[ns] + 1));
  int *fill = (int *)calloc((size_t)ns + 1, sizeof(int));
  for (int c = 0; c < AbNp1->stcf.n; c++) { int s = AbNp1->stcf.a[c]; child[cbeg[s] + fill[s]++] = c; }
  mat_clear(RhoN);
  mat_sync(RhoN, AcN);
  RhoN->upper = 1;
  int i1 = 0;
  double tr = 0.0;
  for (int b = 0; b < nblocks(AcN); b++) {
    ipush(&RhoN->map, b); ipush(&RhoN->map, b); ipush(&RhoN->mbe, i1);
    int i0 = blim(AcN, b, 0), i9 = blim(AcN, b, 1);
    double Sold = getqn(AcN, b, 1);
    for (int ist = i0; ist <= i9; ist++)
      for (int jst = ist; jst <= i9; jst++) {
        double el = 0.0;
        for (int a = cbeg[ist]; a < cbeg[ist + 1]; a++) {
          int ci = child[a], pos;
          int ti = AbNp1->itype.a[ci], bi = blk_from_st(AbNp1, ci, &pos);
          for (int c = cbeg[jst]; c < cbeg[jst + 1]; c++) {
            int cj = child[c], pos2;
            int tj = AbNp1->itype.a[cj], bj = blk_from_st(AbNp1, cj, &pos2);
            if (ti != tj || bi != bj) continue;
            double Si = getqn(AbNp1, bi, 1);
            double St = qn_from_st(S, ti, 1);
            double cg = 0.0;
            if (AcN->sz) cg = 1.0;   /* Sz blocks: the plain partial trace */
            else
              for (double Sz = -Si; Sz <= Si; Sz += 1.0) {
                double x = orc_cgordan(Sold, Sold, St, Sz - Sold, Si, Sz);
                cg += x * x;
              }
            el += cg * rot_el(&rot, rblk, rbeg, ci, cj);
          }
        }
        if (fdm && jst == ist && check_kept(AcN, ist, 0))
          if (fabs(ZNdisc - 0.0) > 1e-20) el = exp(-(1.0 / TempBar) * AcN->en.a[ist]) / ZNdisc;
        if (ist == jst) tr += blk_deg(AcN, b) * el;
        dpush(&RhoN->mat, el); i1++;
      }
    ipush(&RhoN->mbe, i1 - 1);
  }
  free(ccount); free(cbeg); free(child); free(fill); free(rblk); free(rbeg);
  return tr;
}

/* CG factor of a spin-1/2 operator between blocks of spin S1 and S2
 * (src/SpecFunc_allSTL.cpp:654-671, src/SpecFuncClass.cpp:621-634) */
static double op_cgfactor(double S1, double S2) {
  double cg = 0.0;
  for (double Sz = -S1; Sz <= S1; Sz += 1.0) { double x = orc_cgordan(S1, Sz, 0.5, 0.5, S2, Sz + 0.5); cg += x * x; }
  return cg;
}

/* MijxBDeltaEij  src/SpecFunc_allSTL.cpp:36-219: Mout = Min o BD(omega, Ej-Ei), with
 * BDeltaTemp(dBroadTemp) instead of BDelta for |Ej-Ei| < twindow*T/D_N.  BDelta /
 * BDeltaTemp by method (src/DM_NRG/DM_NRG.cpp:474-482): BroadDelta/LorentzDelta
 * (DM-NRG), BroadDelta/LorentzDeltaAnders (CFS), LogGaussDelta/GaussDelta (FDM) */
static double bdelta_ij(const struct orc_dm *D, const arr_t *A, double w, double Ei, double Ej) {
  double TempDN = D->temp / calc_dn(D->p.Lambda, A->nshell);
  double x = Ej - Ei;
  if (fabs(x) < D->p.twindow * TempDN) {
    if (D->method == 1) return lorentz_delta_anders(w, x, D->p.broadtemp);
    if (D->method == 2) return gauss_delta(w, x, D->p.broadtemp);
    return lorentz_delta(w, x, D->p.broadtemp);
  }
  if (D->method == 2) return loggauss_delta(w, x, D->p.bbroad);
  return broad_delta(w, x, D->p.bbroad);
}

/* CalcSpecCosti_T_N  src/SpecFuncClass.cpp:417-662 (real operators); CalcNorm: BD -> 1 */
static double spec_costi(const struct orc_dm *D, int N, double bb, double w, int norm, double zdiv) {
  const arr_t *A = &D->acut[N];
  const mat_t *Op = &D->op[N];
  double acc = 0.0;
  for (int im = 0; im < Op->map.n / 2; im++) {
    int b1 = Op->map.a[2 * im], b2 = Op->map.a[2 * im + 1];
    int n1 = bsize(A, b1), n2 = bsize(A, b2), i0 = blim(A, b1, 0), j0 = blim(A, b2, 0);
    /* Sum12 = opB Rho1 + Rho2 opB (n2 x n1), opB = opA^T; trace of Sum12 opAw */
    double tr = 0.0;
    for (int j = 0; j < n2; j++)
      for (int i = 0; i < n1; i++) {
        double a = get_block_matel(Op, b1, b2, i, j);
        double s12 = a * exp(-bb * A->en.a[i0 + i]) + exp(-bb * A->en.a[j0 + j]) * a;
        tr += s12 * a * (norm ? 1.0 : bdelta_ij(D, A, w, A->en.a[i0 + i], A->en.a[j0 + j]));
      }
    acc += (A->sz ? 1.0 : op_cgfactor(getqn(A, b1, 1), getqn(A, b2, 1))) * tr;
  }
  return acc / (zdiv > 0.0 ? zdiv : partition_func(A, bb));
}

/* DMNRG_SpecDens_M_STL  src/SpecFunc_allSTL.cpp:223-692 (real operators,
 * diagonal GF): per op block pair Termo1 = Tr[A^T rho1 (A o BD)],
 * Termo2 = Tr[rho2 A^T (A o BD)], weighted by the CG factor; the last shell
 * NshellMax-1 uses the Costi form with the input betabar */
static double spec_kd(const struct orc_dm *D, int N, double w, int norm);
double orc_dm_spec_shell(const orc_dm *D, int N, double w) {
  if (N > D->spec_nshell_max - 1) return 0.0;
  if (N == D->spec_nshell_max - 1) {
    /* FDM without the Costi quirk: the last shell's states are all discarded,
     * rho_Nmax = e^{-bE}/Z_tot at the scale of its FDM weight */
    if (D->fix & ORC_FDM_FIX_LAST) return spec_costi(D, N, D->bb_last, w, 0, D->ztot);
    double bb = (fabs(D->temp) < 1e-20) ? 1.0e10 : D->betabar_spec;
    return spec_costi(D, N, bb, w, 0, 0.0);
  }
  if (D->method != 0) return spec_kd(D, N, w, 0);
  const arr_t *A = &D->acut[N];
  const mat_t *Op = &D->op[N], *R = &D->rho[N];
  double acc = 0.0;
  for (int im = 0; im < Op->map.n / 2; im++) {
    int b1 = Op->map.a[2 * im], b2 = Op->map.a[2 * im + 1];
    int n1 = bsize(A, b1), n2 = bsize(A, b2), i0 = blim(A, b1, 0), j0 = blim(A, b2, 0);
    double *Aw = (double *)malloc(sizeof(double) * (size_t)(n1 * n2 + 1));
    double *P = (double *)malloc(sizeof(double) * (size_t)(n1 * n2 + 1));
    /* MatBlock2STL of the operator block and of the two rho blocks */
    double *Ab = (double *)malloc(sizeof(double) * (size_t)(n1 * n2 + 1));
    double *R1 = (double *)malloc(sizeof(double) * (size_t)(n1 * n1 + 1));
    double *R2 = (double *)malloc(sizeof(double) * (size_t)(n2 * n2 + 1));
    for (int i = 0; i < n1; i++) for (int j = 0; j < n2; j++) Ab[i * n2 + j] = get_block_matel(Op, b1, b2, i, j);
    for (int i = 0; i < n1; i++) for (int j = 0; j < n1; j++) R1[i * n1 + j] = get_block_matel(R, b1, b1, i, j);
    for (int i = 0; i < n2; i++) for (int j = 0; j < n2; j++) R2[i * n2 + j] = get_block_matel(R, b2, b2, i, j);
    for (int i = 0; i < n1; i++)
      for (int j = 0; j < n2; j++)
        Aw[i * n2 + j] = Ab[i * n2 + j] * bdelta_ij(D, A, w, A->en.a[i0 + i], A->en.a[j0 + j]);
    /* Termo1: Prod1 = A^T rho1 (n2 x n1); trace(Prod1 Aw) */
    for (int j = 0; j < n2; j++)
      for (int i = 0; i < n1; i++) {
        double s = 0.0;
        for (int l = 0; l < n1; l++) s += Ab[l * n2 + j] * R1[l * n1 + i];
        P[j * n1 + i] = s;
      }
    double t1 = 0.0;
    for (int j = 0; j < n2; j++) for (int i = 0; i < n1; i++) t1 += P[j * n1 + i] * Aw[i * n2 + j];
    /* Termo2: Prod2 = rho2 A^T (n2 x n1); trace(Prod2 Aw) */
    for (int j = 0; j < n2; j++)
      for (int i = 0; i < n1; i++) {
        double s = 0.0;
        for (int l = 0; l < n2; l++) s += R2[j * n2 + l] * Ab[i * n2 + l];
        P[j * n1 + i] = s;
      }
    double t2 = 0.0;
    for (int j = 0; j < n2; j++) for (int i = 0; i < n1; i++) t2 += P[j * n1 + i] * Aw[i * n2 + j];
    acc += (A->sz ? 1.0 : op_cgfactor(getqn(A, b1, 1), getqn(A, b2, 1))) * (t1 + t2);
    free(Aw); free(P); free(Ab); free(R1); free(R2);
  }
  return acc;
}

/* Terms of the kept/discarded split, on index sets of the two blocks (positions
 * inside the blocks); A is the n1 x n2 operator block, Dl the n1 x n2 broadening.
 *   rho1 side (Prod1 = B^T rho1, trace with A o BD):
 *     sum_{j in J} sum_{i in I} (sum_{l in I} A_lj R1_li) A_ij Dl_ij
 *   rho2 side (Prod2 = rho2 B^T):
 *     sum_{j in J} sum_{i in I} (sum_{l in J} R2_jl A_il) A_ij Dl_ij */
static double kd_term1(const double *A, int n2, const double *R1, int n1, const int *I, int ni, const int *J, int nj,
                       const double *Dl) {
  double t = 0.0;
  for (int jj = 0; jj < nj; jj++) {
    int j = J[jj];
    for (int ii = 0; ii < ni; ii++) {
      int i = I[ii];
      double p = 0.0;
      for (int ll = 0; ll < ni; ll++) p += A[(size_t)I[ll] * n2 + j] * R1[(size_t)I[ll] * n1 + i];
      t += p * A[(size_t)i * n2 + j] * Dl[(size_t)i * n2 + j];
    }
  }
  return t;
}
static double kd_term2(const double *A, int n2, const double *R2, const int *I, int ni, const int *J, int nj,
                       const double *Dl) {
  double t = 0.0;
  for (int jj = 0; jj < nj; jj++) {
    int j = J[jj];
    for (int ii = 0; ii < ni; ii++) {
      int i = I[ii];
      double p = 0.0;
      for (int ll = 0; ll < nj; ll++) p += R2[(size_t)j * n2 + J[ll]] * A[(size_t)i * n2 + J[ll]];
      t += p * A[(size_t)i * n2 + j] * Dl[(size_t)i * n2 + j];
    }
  }
  return t;
}

/* CFS_SpecDens_M_STL  src/SpecFunc_allSTL.cpp:702-1240 (method 1) and
 * FDM_SpecDens_M_STL  src/SpecFunc_allSTL.cpp:1253-2354 (method 2), real
 * operators, diagonal GF, below the last shell.  Per operator block pair
 * (b1 -> b2) with K/D = kept/discarded states (MatBlock2STL with kept flags,
 * src/NRGmatrixClass.cpp:1188-1212):
 *   the pair is skipped when neither block has a kept state or neither has a
 *   discarded one;
 *   Termo1KK  = T1(K1, D2, rho1_KK)   if K1, D2 and (w >= 0 or CalcNorm)
 *   Termo2KK  = T2(D1, K2, rho2_KK)   if D1, K2 and (w <  0 or CalcNorm)
 *   FDM only, if D1 and D2:  Termo1DDADD = T1(D1, D2, rho1_DD),
 *     Termo1DDADK = T1(D1, K2, rho1_DD) if K2, Termo2DDADD = T2(D1, D2, rho2_DD),
 *     Termo2DDAKD = T2(K1, D2, rho2_DD) if K1;
 *   Positive += CG (T1KK + T1DDADD + T2DDADD + T2DDAKD),
 *   Negative += CG (T2KK + T1DDADD + T1DDADK + T2DDADD)   (CFS: T1KK, T2KK only);
 * result Positive (w >= 0) or Negative (w < 0); CalcNorm: BD -> 1 and
 * Positive + Negative. */
static double spec_kd(const struct orc_dm *D, int N, double w, int norm) {
  const arr_t *A = &D->acut[N];
  const mat_t *Op = &D->op[N], *R = &D->rho[N];
  const int fdm = (D->method == 2);
  double pos = 0.0, neg = 0.0;
  for (int im = 0; im < Op->map.n / 2; im++) {
    int b1 = Op->map.a[2 * im], b2 = Op->map.a[2 * im + 1];
    int n1 = bsize(A, b1), n2 = bsize(A, b2), i0 = blim(A, b1, 0), j0 = blim(A, b2, 0);
    int k1 = bsize_kp(A, b1, 1), d1 = bsize_kp(A, b1, 0), k2 = bsize_kp(A, b2, 1), d2 = bsize_kp(A, b2, 0);
    if (n1 != k1 + d1 || n2 != k2 + d2) return 0.0;   /* "Kept+Disc NOT equal to Nstblock" */
    if (d1 == 0 && d2 == 0) continue;   /* kept-kept only: no pole at this shell */
    if (k1 == 0 && k2 == 0 && !(D->fix & ORC_FDM_FIX_PAIRS)) continue;
    int *K1 = (int *)malloc(sizeof(int) * (size_t)(n1 + 1)), *D1 = (int *)malloc(sizeof(int) * (size_t)(n1 + 1));
    int *K2 = (int *)malloc(sizeof(int) * (size_t)(n2 + 1)), *D2 = (int *)malloc(sizeof(int) * (size_t)(n2 + 1));
    int a = 0, b = 0;
    for (int i = 0; i < n1; i++) { if (check_kept(A, i0 + i, 1)) K1[a++] = i; if (check_kept(A, i0 + i, 0)) D1[b++] = i; }
    a = 0; b = 0;
    for (int j = 0; j < n2; j++) { if (check_kept(A, j0 + j, 1)) K2[a++] = j; if (check_kept(A, j0 + j, 0)) D2[b++] = j; }
    double *Ab = (double *)malloc(sizeof(double) * (size_t)(n1 * n2 + 1));
    double *Dl = (double *)malloc(sizeof(double) * (size_t)(n1 * n2 + 1));
    double *R1 = (double *)malloc(sizeof(double) * (size_t)(n1 * n1 + 1));
    double *R2 = (double *)malloc(sizeof(double) * (size_t)(n2 * n2 + 1));
    for (int i = 0; i < n1; i++)
      for (int j = 0; j < n2; j++) {
        Ab[i * n2 + j] = get_block_matel(Op, b1, b2, i, j);
        Dl[i * n2 + j] = norm ? 1.0 : bdelta_ij(D, A, w, A->en.a[i0 + i], A->en.a[j0 + j]);
      }
    for (int i = 0; i < n1; i++) for (int j = 0; j < n1; j++) R1[i * n1 + j] = get_block_matel(R, b1, b1, i, j);
    for (int i = 0; i < n2; i++) for (int j = 0; j < n2; j++) R2[i * n2 + j] = get_block_matel(R, b2, b2, i, j);
    /* ORC_FDM_FIX_NORM: every term at both signs of omega (the broadening picks the side) */
    const int both = fdm && (D->fix & ORC_FDM_FIX_NORM);
    const int onesided = fdm && (D->fix & ORC_FDM_FIX_ONESIDED);
    double t1kk = 0, t2kk = 0, t1dd = 0, t1dk = 0, t2dd = 0, t2kd = 0;
    if (k1 && d2 && (w >= 0.0 || norm || both)) t1kk = kd_term1(Ab, n2, R1, n1, K1, k1, D2, d2, Dl);
    if (fdm && d1 && d2) t1dd = kd_term1(Ab, n2, R1, n1, D1, d1, D2, d2, Dl);
    if (fdm && d1 && k2 && (d2 || onesided)) t1dk = kd_term1(Ab, n2, R1, n1, D1, d1, K2, k2, Dl);
    if (d1 && k2 && (w < 0.0 || norm || both)) t2kk = kd_term2(Ab, n2, R2, D1, d1, K2, k2, Dl);
    if (fdm && d1 && d2) t2dd = kd_term2(Ab, n2, R2, D1, d1, D2, d2, Dl);
    if (fdm && k1 && d2 && (d1 || onesided)) t2kd = kd_term2(Ab, n2, R2, K1, k1, D2, d2, Dl);
    double cg = A->sz ? 1.0 : op_cgfactor(getqn(A, b1, 1), getqn(A, b2, 1));
    if (both) {
      pos += cg * (t1kk + t2kk + t1dd + t1dk + t2dd + t2kd);
    } else {
      pos += cg * (t1kk + t1dd + t2dd + t2kd);
      neg += cg * (t2kk + t1dd + t1dk + t2dd);
    }
    free(K1); free(D1); free(K2); free(D2); free(Ab); free(Dl); free(R1); free(R2);
  }
  if (fdm && (D->fix & ORC_FDM_FIX_NORM)) return pos;
  if (norm) return pos + neg;
  return (w >= 0.0) ? pos : neg;
}

/* CalcNfromOmega  src/SpecFuncClass.cpp:41-49 (z = 1) */
static int n_from_omega(double L, double w) {
  double HLF = 0.5 * (1.0 + 1.0 / L);
  int i = (int)(2.0 * log(2.0 * fabs(w) / HLF) / log(L));
  i = 1 - i;
  return i < 0 ? 0 : i;
}
double orc_dm_spec_omega(const orc_dm *D, double w) {
  int No = n_from_omega(D->p.Lambda, w);
  int N0 = (No - 8 > 0) ? No - 8 : 0;
  int N1 = (No + 14 < D->spec_nshell_max) ? No + 14 : D->spec_nshell_max;
  double s = 0.0;
  for (int N = N0; N < N1; N++) {
    double DN = calc_dn(D->p.Lambda, N);
    s += orc_dm_spec_shell(D, N, w / DN) / DN;
  }
  return s;
}
double orc_dm_norm(const orc_dm *D) {
  if (D->method == 0) return NAN;
  double s = 0.0;
  for (int N = 0; N < D->spec_nshell_max; N++) {
    if (N == D->spec_nshell_max - 1) {
      if (D->fix & ORC_FDM_FIX_LAST) { s += spec_costi(D, N, D->bb_last, 0.0, 1, D->ztot); continue; }
      double bb = (fabs(D->temp) < 1e-20) ? 1.0e10 : D->betabar_spec;
      s += spec_costi(D, N, bb, 0.0, 1, 0.0);
    } else {
      s += spec_kd(D, N, 0.0, 1);
    }
  }
  return s;
}
double orc_dm_ztot(const orc_dm *D) { return D->ztot; }

static void dm_store_step(struct orc_dm *D, const arr_t *Ac, const arr_t *Ab, const mat_t *op) {
  int N = D->nsh++;
  D->acut = (arr_t *)realloc(D->acut, sizeof(arr_t) * D->nsh);
  D->abasis = (arr_t *)realloc(D->abasis, sizeof(arr_t) * D->nsh);
  D->op = (mat_t *)realloc(D->op, sizeof(mat_t) * D->nsh);
  memset(&D->acut[N], 0, sizeof(arr_t)); memset(&D->abasis[N], 0, sizeof(arr_t)); memset(&D->op[N], 0, sizeof(mat_t));
  arr_copy(&D->acut[N], Ac); arr_copy(&D->abasis[N], Ab); mat_copy(&D->op[N], op);
}

/* gsl_interp_linear evaluation on ascending knots (GSL interp/linear.c) */
static double lin_interp(const double *x, const double *y, int n, double t) {
  int lo = 0, hi = n - 1;
  while (hi > lo + 1) { int m = (lo + hi) / 2; if (x[m] > t) hi = m; else lo = m; }
  double dx = x[hi] - x[lo];
  if (dx > 0.0) return y[lo] + (t - x[lo]) * ((y[hi] - y[lo]) / dx);
  return y[lo];
}
static int cmpd2(const void *a, const void *b) {
  double x = *(const double *)a, y = *(const double *)b;
  return (x > y) - (x < y);
}

/* The evaluation points of the omega-grid drivers, in the order the reference
 * records them: shell N, the omegabar its DM-NRG shell sum is evaluated at, the
 * recorded omega, and the even/odd set (FixedOmegas; -1 for ManyOmegas).
 *   CalcSpecDM_NRG_FixedOmegas  src/SpecFuncClass.cpp:1932-2233 (NwEachN = 1, the call of
 *     DM_NRG_CalcSpecFunc_ij src/DM_NRG/DM_NRG_CalcSpecFuncs.cpp:83-86): even shells from
 *     Nsh0Even, w = -Lambda^1.25 D_N ascending in N, then +w back down; the same for the odd
 *     shells; windowed (CFS/FDM) drivers start two shells later (:1954);
 *   CalcSpec_ManyOmegas  :2239-2341: shells Nsh0 .. NshellMax-1 (+2 windowed), recorded
 *     w = -+Lambda^(-i/(2Nw)) Lambda^1.25 D_N, i = 0 .. Nw-1 (negative: N up, i up; positive:
 *     N down, i down); the DM-NRG shell sum is evaluated at -+Lambda^1.25 Lambda^(+i/(2Nw)).
 * Returns the count; the arrays may be NULL to query it. */
int orc_grid_plan(double L, int nshell, int windowed, int Nw, int *shell, double *wbar, double *omega,
                  int *parity) {
  const double fac = pow(L, 1.25);
  int k = 0;
#define ORC_GP(N_, wb_, w_, par_)                                    \
  do {                                                               \
    if (shell) { shell[k] = (N_); wbar[k] = (wb_); omega[k] = (w_); parity[k] = (par_); } \
    k++;                                                             \
  } while (0)
  if (Nw > 0) {
    int N0 = windowed ? 2 : 0, Nmax = nshell - 1;
    for (int N = N0; N <= Nmax; N++)
      for (int i = 0; i < Nw; i++) {
        double rf = (double)i / (2.0 * Nw);
        ORC_GP(N, -fac * pow(L, rf), -pow(L, -rf) * fac * calc_dn(L, N), -1);
      }
    for (int N = Nmax; N >= N0; N--)
      for (int i = Nw - 1; i >= 0; i--) {
        double rf = (double)i / (2.0 * Nw);
        ORC_GP(N, fac * pow(L, rf), pow(L, -rf) * fac * calc_dn(L, N), -1);
      }
  } else {
    for (int par = 0; par < 2; par++) {
      int n0 = par + (windowed ? 2 : 0), last = 0;
      for (int N = n0; N < nshell; N += 2) { ORC_GP(N, -fac, -fac * calc_dn(L, N), par); last = N; }
      for (int N = last; N >= n0; N -= 2) ORC_GP(N, fac, fac * calc_dn(L, N), par);
    }
  }
#undef ORC_GP
  return k;
}

/* The spectrum from the values at the plan's points: ManyOmegas records them
 * as they are; FixedOmegas interpolates the even and the odd set linearly
 * (gsl_interp_linear) on the sorted union of their omegas and averages, a
 * point outside one set's range taking the other's value (outside both: -100)
 * (src/SpecFuncClass.cpp:2112-2222).  Returns the count of (w_out, rho_out). */
int orc_grid_combine(int n, const double *omega, const int *parity, const double *val, int Nw, double *w_out,
                     double *rho_out) {
  if (Nw > 0) {
    for (int i = 0; i < n; i++) { w_out[i] = omega[i]; rho_out[i] = val[i]; }
    return n;
  }
  dvec ex = {0}, ey = {0}, ox = {0}, oy = {0};
  for (int i = 0; i < n; i++) {
    if (parity[i] == 0) { dpush(&ex, omega[i]); dpush(&ey, val[i]); }
    else { dpush(&ox, omega[i]); dpush(&oy, val[i]); }
    w_out[i] = omega[i];
  }
  qsort(w_out, (size_t)n, sizeof(double), cmpd2);
  for (int i = 0; i < n; i++) {
    double w = w_out[i];
    int inE = ex.n && !(w < ex.a[0] || w > ex.a[ex.n - 1]), inO = ox.n && !(w < ox.a[0] || w > ox.a[ox.n - 1]);
    double rE = inE ? lin_interp(ex.a, ey.a, ex.n, w) : 0.0, rO = inO ? lin_interp(ox.a, oy.a, ox.n, w) : 0.0;
    if (!inE) { if (!inO) { rE = -100.0; rO = -100.0; } else rE = rO; }
    else if (!inO) rO = rE;
    rho_out[i] = 0.5 * (rE + rO);
  }
  free(ex.a); free(ey.a); free(ox.a); free(oy.a);
  return n;
}

/* rho_0_0 on the grid: DM-NRG evaluates each shell's own sum (/D_N); CFS and
 * FDM call CalcSpecDM_NRG at the recorded omega (src/SpecFuncClass.cpp:1979-1982) */
static void spec_grid(struct orc_dm *D) {
  const int win = D->method != 0;
  int n = orc_grid_plan(D->p.Lambda, D->spec_nshell_max, win, D->p.Nw, NULL, NULL, NULL, NULL);
  int *sh = (int *)malloc(sizeof(int) * (size_t)(n + 1)), *par = (int *)malloc(sizeof(int) * (size_t)(n + 1));
  double *wb = (double *)malloc(sizeof(double) * (size_t)(n + 1)), *w = (double *)malloc(sizeof(double) * (size_t)(n + 1));
  double *v = (double *)malloc(sizeof(double) * (size_t)(n + 1));
  orc_grid_plan(D->p.Lambda, D->spec_nshell_max, win, D->p.Nw, sh, wb, w, par);
  for (int i = 0; i < n; i++)
    v[i] = win ? orc_dm_spec_omega(D, w[i]) : orc_dm_spec_shell(D, sh[i], wb[i]) / calc_dn(D->p.Lambda, sh[i]);
  D->omega = (double *)malloc(sizeof(double) * (size_t)(n + 1));
  D->rhow = (double *)malloc(sizeof(double) * (size_t)(n + 1));
  D->nomega = orc_grid_combine(n, w, par, v, D->p.Nw, D->omega, D->rhow);
  free(sh); free(par); free(wb); free(w); free(v);
}

/* DM_NRG main  src/DM_NRG/DM_NRG.cpp:28-531 (QS symmetry, NFermiOps 1):
 * temperature from Mtemp (:102-118; FDM keeps every shell, :105), rho at
 * NshellMax with ZN = PartitionFunc at Temp/D_{NshellMax-1} (:203-214, the
 * reference's own "Why????" shell offset kept); FDM: w_N from Z_N^disc 4^{Nmax-N}
 * (:221-271) and rho_Nmax normalised by Z_tot with the exponent Temp/DM of
 * :102-118 (:273-297); backward sweep (:312-434) with TempBar = Temp/D_{N-1} and
 * ZNdisc = Z_tot/4^{Nmax-N} (:361-366); spectral function rho_0_0 (:451-506) on the grid
 * of DM_NRG_CalcSpecFunc_ij src/DM_NRG/DM_NRG_CalcSpecFuncs.cpp:73-84 */
static void tan_dm_sweep(struct orc_dm *D, double bb_r, double bb_z, double ZN);
/* a TwoChQSz shell as SaveArrays would write it with UpdateBefCut 1: every state
 * of Aeig with its Kept flags, Abasis with StCameFrom as the previous shell's
 * indices before its cut, d_up on every state */
static void dm_store_qsz(struct orc_dm *D, const arr_t *Ae, const arr_t *Ab, const arr_t *Acprev, const mat_t *dall) {
  arr_t A = {0}, B = {0};
  arr_copy(&A, Ae);
  A.sz = 1;
  icopy_v(&A.bbe_bc, &Ae->bbe);
  A.stcf.n = 0;
  for (int i = 0; i < A.en.n; i++) ipush(&A.stcf, i);
  set_bbe_ev(&A);
  arr_copy(&B, Ab);
  B.sz = 1;
  for (int i = 0; i < B.stcf.n; i++) B.stcf.a[i] = Acprev->stcf.a[Ab->stcf.a[i]];
  mat_t O = {0};
  mat_copy(&O, dall);
  O.b.sz = 1;
  dm_store_step(D, &A, &B, &O);
  mat_free(&O);
  arr_free(&A); arr_free(&B);
}
orc_dm *orc_dm_run(const orc_dmnrg_params *p) { return orc_dm_run_jvp(p, NULL, NULL, 0, NULL); }
orc_dm *orc_dm_run_jvp(const orc_dmnrg_params *p, const unsigned char *kept, const long long *kept_off, int ntan,
                       const double *dpar) {
  if (ntan < 0 || ntan > ORC_MAX_TAN || (ntan > 0 && (!dpar || p->method != 0 || p->update_bef_cut || !(p->Gamma > 0))))
    return NULL;
  if (p->model == 1 && (ntan > 0 || kept)) return NULL;
  struct orc_dm *D = (struct orc_dm *)calloc(1, sizeof(struct orc_dm));
  D->p = *p;
  if (D->p.twindow <= 0) D->p.twindow = 2.0;
  if (D->p.broadtemp <= 0) D->p.broadtemp = 0.8;
  if (D->p.bbroad <= 0) D->p.bbroad = 0.5 * log(p->Lambda);
  D->method = (p->method == 1 || p->method == 2) ? p->method : 0;
  const int fdm = (D->method == 2);
  orc_onechqs_params q;
  memset(&q, 0, sizeof(q));
  q.U = p->U; q.Gamma = p->Gamma; q.ed = p->ed; q.Lambda = p->Lambda; q.nsitesmax = p->nsitesmax;
  q.ncut = p->ncut; q.calcdens = 1; q.disc_scheme = p->disc_scheme; q.myeps = p->myeps;
  q.betabar = p->betabar_nrg;
  orc_tan T;
  memset(&T, 0, sizeof(T));
  T.ntan = ntan;
  for (int t = 0; t < ntan; t++)
    for (int c = 0; c < 3; c++) T.dpar[t][c] = dpar[3 * t + c];
  D->ntan = ntan;
  const int qsz = (p->model == 1);
  if (qsz) {   /* TwoChQSz standalone loop (src/TwoChQSz/TwoChQSz.cpp:103-425), shells 0 .. nsitesmax */
    orc_twochqs_params q2;
    memset(&q2, 0, sizeof(q2));
    q2.U = p->U; q2.Gamma1 = p->Gamma; q2.Gamma2 = p->Gamma2; q2.ed = p->ed; q2.Lambda = p->Lambda;
    q2.nsitesmax = p->nsitesmax; q2.ncut = p->ncut; q2.betabar = p->betabar_nrg;
    run_twochqsz(&q2, NULL, NULL, D);
    twochqsz_site(&D->site);
  } else {
    run_onechqs(&q, NULL, NULL, D, p->update_bef_cut ? 1 : 0, kept, kept_off, ntan > 0 ? &T : NULL);
    onechqs_site(&D->site);
  }
  tan_free(&T);
  /* 4: the reference's one-channel value (src/DM_NRG/DM_NRG.cpp:244); 16 the two-channel site */
  const double site_dim = p->fdm_site_dim > 0.0 ? p->fdm_site_dim : (qsz ? 16.0 : 4.0);
  int Nsitesmax = D->nsh;
  double DM, Temp;
  if (p->Mtemp < Nsitesmax - 1) {
    if (!fdm) Nsitesmax = p->Mtemp + 1;
    DM = calc_dn(p->Lambda, p->Mtemp);
    Temp = DM / p->betabar;
  } else {
    Temp = calc_dn(p->Lambda, Nsitesmax + 199) / p->betabar;
    DM = calc_dn(p->Lambda, Nsitesmax - 1);
  }
  D->temp = Temp;
  D->betabar_spec = p->betabar;
  D->nshell_max = Nsitesmax - 1;
  D->spec_nshell_max = Nsitesmax;
  D->fix = fdm ? (p->fdm_fix & ORC_FDM_EXACT) : 0;
  /* shell N's Boltzmann factors: Temp/D_{N-1} as the reference (src/DM_NRG/DM_NRG.cpp:203-214,240,349),
   * or Temp/D_N without that quirk */
  const int dsh = (D->fix & ORC_FDM_FIX_DN) ? 0 : 1;
  int NM = D->nshell_max;
  D->rho = (mat_t *)calloc((size_t)NM + 1, sizeof(mat_t));
  D->trace = (double *)calloc((size_t)NM + 1, sizeof(double));
  double TempBar = Temp / calc_dn(p->Lambda, NM - dsh);
  double ZN = partition_func(&D->acut[NM], 1.0 / TempBar);
  D->wn = (double *)calloc((size_t)NM + 1, sizeof(double));
  double ZTot = ZN;
  D->wn[NM] = ZN;
  for (int N = NM - 1; N >= 0; N--) {
    double tb = Temp / calc_dn(p->Lambda, N - dsh);
    double zd = partition_func_disc(&D->acut[N], 1.0 / tb) * pow(site_dim, NM - N);
    D->wn[N] = zd;
    ZTot += zd;
  }
  for (int N = 0; N <= NM; N++) D->wn[N] = fdm ? D->wn[N] / ZTot : 0.0;
  D->ztot = fdm ? ZTot : 0.0;
  D->bb_last = 1.0 / TempBar;
  /* rho_Nmax's exponent: Temp/D_Mtemp as the reference (src/DM_NRG/DM_NRG.cpp:102-118,273-297), or the
   * scale of Z_tot and w_Nmax */
  D->trace[NM] = set_rho_nmax((D->fix & ORC_FDM_FIX_RHOMAX) ? TempBar : Temp / DM, fdm ? ZTot : ZN, &D->acut[NM],
                              &D->rho[NM]);
  for (int N = NM - 1; N >= 0; N--) {
    double tb = Temp / calc_dn(p->Lambda, N - dsh);
    D->trace[N] = calc_rho_n(&D->acut[N], &D->acut[N + 1], &D->abasis[N + 1], &D->site, &D->rho[N + 1], &D->rho[N],
                             fdm, tb, ZTot / pow(site_dim, NM - N));
  }
  if (ntan > 0) tan_dm_sweep(D, fabs(Temp / DM) < 1e-20 ? 1e10 : DM / Temp, 1.0 / TempBar, ZN);
  return D;
}

/* ------------------------------------------------------------------ */
/* Derivative of the DM-NRG shell sums (method 0) along the tangents of  */
/* the fixed-truncation chain; no counterpart in the reference.         */
/* The chain's tangents live in the kept-space frame (kept rows move    */
/* only towards discarded ones; the kept-space Hamiltonian tangent th   */
/* is a full block).  rho_Nmax = exp(-b H)/Z_N and the backward sweep   */
/* rho_N = Tr_site[Z^T rho_{N+1} Z] are differentiated in that frame;   */
/* a shell sum is then taken to the perturbed eigenbasis by            */
/* M_ij = th_ij / (E_i - E_j) for non-degenerate pairs (dX -> dX +      */
/* M1 X - X M2), while inside degenerate clusters th stays and moves    */
/* the poles: with W_ij = cg A_ij Q_ij, Q = rho1 A + A rho2,            */
/*   dA_N(w) = sum_ij W1_ij delta(w, E_j - E_i) + W2_ij d delta/dE,     */
/*   W1 = cg (dA Q + A dQ), W2 = cg Q o (A th2_cl - th1_cl A).         */
/* ------------------------------------------------------------------ */
#define ORC_DEG_TOL 1e-9
static void dm_store_tan(struct orc_dm *D, const orc_tan *T) {
  int N = D->nsh - 1, nsh = D->nsh;
  D->th = realloc(D->th, sizeof(*D->th) * (size_t)nsh);
  D->th_off = (ivec *)realloc(D->th_off, sizeof(ivec) * (size_t)nsh);
  D->dop = realloc(D->dop, sizeof(*D->dop) * (size_t)nsh);
  D->dz = realloc(D->dz, sizeof(*D->dz) * (size_t)nsh);
  memset(&D->th[N], 0, sizeof(D->th[N])); memset(&D->th_off[N], 0, sizeof(ivec));
  memset(&D->dop[N], 0, sizeof(D->dop[N])); memset(&D->dz[N], 0, sizeof(D->dz[N]));
  icopy_v(&D->th_off[N], &T->th_off);
  for (int t = 0; t < T->ntan; t++) {
    dcopy_v(&D->th[N][t], &T->th[t]);
    dcopy_v(&D->dop[N][t], &T->tm[t][1]);
    dcopy_v(&D->dz[N][t], &T->dz[t]);
  }
}
/* element (i, j) of block pair (b1, b2) of a tangent stored in M's layout */
static double get_block_tan(const mat_t *M, const dvec *tv, int b1, int b2, int i, int j) {
  int im = find_mat_block(M, b1, b2);
  if (im < 0) return 0.0;
  int n1 = bsize(&M->b, b1), n2 = bsize(&M->b, b2), r = n2 * i + j;
  if (M->upper && b1 == b2) r = (i > j) ? ij2rUpTr(n1, j, i) : ij2rUpTr(n1, i, j);
  return tv->a[M->mbe.a[2 * im] + r];
}
/* divided difference of exp(-b E) between Ei and Ej */
static double dd_exp(double b, double Ei, double Ej) {
  double d = Ei - Ej, xj = exp(-b * Ej);
  return d != 0.0 ? xj * expm1(-b * d) / d : -b * xj;
}
static void tan_dm_sweep(struct orc_dm *D, double bb_r, double bb_z, double ZN) {
  int NM = D->nshell_max;
  D->drho = calloc((size_t)NM + 1, sizeof(*D->drho));
  const arr_t *A = &D->acut[NM];
  for (int t = 0; t < D->ntan; t++) {
    /* rho_Nmax = exp(-bb_r H) / Z_N, Z_N = sum (2S+1) exp(-bb_z E) (the reference's two shells) */
    const double *th = D->th[NM][t].a;
    const int *off = D->th_off[NM].a;
    double dZ = 0.0;
    for (int b = 0; b < nblocks(A); b++) {
      int k = bsize(A, b), b0 = blim(A, b, 0);
      double dg = 2.0 * getqn(A, b, 1) + 1.0;
      for (int i = 0; i < k; i++) dZ += dg * th[off[b] + i * k + i] * (-bb_z * exp(-bb_z * A->en.a[b0 + i]));
    }
    dvec *out = &D->drho[NM][t];
    for (int b = 0; b < nblocks(A); b++) {   /* packed upper, the layout of set_rho_nmax */
      int k = bsize(A, b), b0 = blim(A, b, 0);
      for (int i = 0; i < k; i++)
        for (int j = i; j < k; j++) {
          double Ei = A->en.a[b0 + i], Ej = A->en.a[b0 + j];
          double rho = (i == j) ? exp(-bb_r * Ei) / ZN : 0.0;
          dpush(out, dd_exp(bb_r, Ei, Ej) * th[off[b] + i * k + j] / ZN - rho * dZ / ZN);
        }
    }
  }
  for (int N = NM - 1; N >= 0; N--) {
    const arr_t *Ac = &D->acut[N + 1];
    const mat_t *R = &D->rho[N + 1];
    for (int t = 0; t < D->ntan; t++) {
      /* Z^T drho Z + dZ^T rho Z + Z^T rho dZ over the kept rows of shell N+1 */
      mat_t tmp = {0}, rot = {0}, g = {0};
      mat_copy(&tmp, R);
      dcopy_v(&tmp.mat, &D->drho[N + 1][t]);
      rotate_matrix(&tmp, Ac, &rot);
      const double *dzb = D->dz[N + 1][t].a;
      int i1 = 0;
      for (int im = 0; im < R->map.n / 2; im++) {
        int b1 = R->map.a[2 * im], b2 = R->map.a[2 * im + 1];
        int k1 = bsize(Ac, b1), k2 = bsize(Ac, b2), n1 = bsize_bc(Ac, b1), n2 = bsize_bc(Ac, b2);
        const double *Z1 = Ac->ev.a + Ac->bbe_ev.a[2 * b1], *Z2 = Ac->ev.a + Ac->bbe_ev.a[2 * b2];
        const double *dZ1 = dzb + Ac->bbe_ev.a[2 * b1], *dZ2 = dzb + Ac->bbe_ev.a[2 * b2];
        double *MZ = (double *)calloc((size_t)k1 * n2 + 1, sizeof(double));
        double *MdZ = (double *)calloc((size_t)k1 * n2 + 1, sizeof(double));
        for (int i = 0; i < k1; i++)
          for (int j = 0; j < k2; j++) {
            double m = get_block_matel(R, b1, b2, i, j);
            for (int c = 0; c < n2; c++) { MZ[i * n2 + c] += m * Z2[j * n2 + c]; MdZ[i * n2 + c] += m * dZ2[j * n2 + c]; }
          }
        for (int r = 0; r < n1; r++)
          for (int c = (rot.upper && b1 == b2) ? r : 0; c < n2; c++) {
            double v = 0.0;
            for (int i = 0; i < k1; i++) v += dZ1[i * n1 + r] * MZ[i * n2 + c] + Z1[i * n1 + r] * MdZ[i * n2 + c];
            rot.mat.a[i1++] += v;
          }
        free(MZ); free(MdZ);
      }
      gather_rho(&D->acut[N], &D->abasis[N + 1], &D->site, &rot, &g, 0, 1.0, 0.0);
      dcopy_v(&D->drho[N][t], &g.mat);
      mat_free(&tmp); mat_free(&rot); mat_free(&g);
    }
  }
}
/* BDelta / BDeltaTemp of bdelta_ij and their derivatives with respect to the pole energy */
static double bdelta_ij_d(const struct orc_dm *D, const arr_t *A, double w, double Ei, double Ej, double *gp) {
  double TempDN = D->temp / calc_dn(D->p.Lambda, A->nshell);
  double x = Ej - Ei;
  if (fabs(x) < D->p.twindow * TempDN) {
    double b = D->p.broadtemp, u = w - x, q = u * u + b * b;
    *gp = (1.0 / (2.0 * M_PI)) * b * 2.0 * u / (q * q);
    return lorentz_delta(w, x, b);
  }
  double g = broad_delta(w, x, D->p.bbroad);
  double Ep = (x == 0.0) ? 1.0e-100 : x;
  *gp = (w * Ep < 0.0) ? 0.0 : g * (2.0 * log(fabs(w / Ep)) / (D->p.bbroad * D->p.bbroad) - 1.0) / Ep;
  return g;
}
/* d/dtheta_t of orc_dm_spec_shell(D, N, w) for every tangent t */
int orc_dm_spec_shell_jvp(const orc_dm *D, int N, double w, double *out) {
  if (D->ntan <= 0 || N < 0 || N > D->spec_nshell_max - 1) return -1;
  const arr_t *A = &D->acut[N];
  const mat_t *Op = &D->op[N];
  const int costi = (N == D->spec_nshell_max - 1);
  const double bb = (fabs(D->temp) < 1e-20) ? 1.0e10 : D->betabar_spec;
  const double Zc = costi ? partition_func(A, bb) : 1.0;
  for (int t = 0; t < D->ntan; t++) {
    const double *th = D->th[N][t].a;
    const int *off = D->th_off[N].a;
    double dZc = 0.0;
    if (costi)
      for (int b = 0; b < nblocks(A); b++) {
        int k = bsize(A, b), b0 = blim(A, b, 0);
        double dg = 2.0 * getqn(A, b, 1) + 1.0;
        for (int i = 0; i < k; i++) dZc += dg * th[off[b] + i * k + i] * (-bb * exp(-bb * A->en.a[b0 + i]));
      }
    double acc = 0.0;
    for (int im = 0; im < Op->map.n / 2; im++) {
      int bl[2] = {Op->map.a[2 * im], Op->map.a[2 * im + 1]};
      int n[2], e0[2];
      double *R[2], *dR[2], *Mr[2], *Tc[2];
      for (int s = 0; s < 2; s++) {
        int b = bl[s], k = bsize(A, b);
        n[s] = k; e0[s] = blim(A, b, 0);
        R[s] = (double *)calloc((size_t)k * k + 1, sizeof(double));
        dR[s] = (double *)calloc((size_t)k * k + 1, sizeof(double));
        Mr[s] = (double *)calloc((size_t)k * k + 1, sizeof(double));
        Tc[s] = (double *)calloc((size_t)k * k + 1, sizeof(double));
        const double *E = A->en.a + e0[s], *tb = th + off[b];
        for (int i = 0; i < k; i++)
          for (int j = 0; j < k; j++) {
            double d = E[i] - E[j];
            if (fabs(d) > ORC_DEG_TOL) Mr[s][i * k + j] = tb[i * k + j] / d; else Tc[s][i * k + j] = tb[i * k + j];
            if (costi) {   /* rho = exp(-bb H)/Z in the kept-space frame */
              R[s][i * k + j] = (i == j) ? exp(-bb * E[i]) / Zc : 0.0;
              dR[s][i * k + j] = dd_exp(bb, E[i], E[j]) * tb[i * k + j] / Zc - R[s][i * k + j] * dZc / Zc;
            } else {
              R[s][i * k + j] = get_block_matel(&D->rho[N], b, b, i, j);
              dR[s][i * k + j] = get_block_tan(&D->rho[N], &D->drho[N][t], b, b, i, j);
            }
          }
      }
      int n1 = n[0], n2 = n[1];
      double *Ab = (double *)malloc(sizeof(double) * (size_t)(n1 * n2 + 1));
      double *dA = (double *)malloc(sizeof(double) * (size_t)(n1 * n2 + 1));
      for (int i = 0; i < n1; i++)
        for (int j = 0; j < n2; j++) {
          Ab[i * n2 + j] = get_block_matel(Op, bl[0], bl[1], i, j);
          dA[i * n2 + j] = get_block_tan(Op, &D->dop[N][t], bl[0], bl[1], i, j);
        }
      /* to the perturbed eigenbasis: dA += M1 A - A M2, drho += M rho - rho M */
      double *dAe = (double *)malloc(sizeof(double) * (size_t)(n1 * n2 + 1));
      for (int i = 0; i < n1; i++)
        for (int j = 0; j < n2; j++) {
          double v = dA[i * n2 + j];
          for (int l = 0; l < n1; l++) v += Mr[0][i * n1 + l] * Ab[l * n2 + j];
          for (int l = 0; l < n2; l++) v -= Ab[i * n2 + l] * Mr[1][l * n2 + j];
          dAe[i * n2 + j] = v;
        }
      double *dRe[2];
      for (int s = 0; s < 2; s++) {
        int k = n[s];
        dRe[s] = (double *)malloc(sizeof(double) * (size_t)(k * k + 1));
        for (int i = 0; i < k; i++)
          for (int j = 0; j < k; j++) {
            double v = dR[s][i * k + j];
            for (int l = 0; l < k; l++) v += Mr[s][i * k + l] * R[s][l * k + j] - R[s][i * k + l] * Mr[s][l * k + j];
            dRe[s][i * k + j] = v;
          }
      }
      double cg = op_cgfactor(getqn(A, bl[0], 1), getqn(A, bl[1], 1));
      for (int i = 0; i < n1; i++)
        for (int j = 0; j < n2; j++) {
          double Q = 0.0, dQ = 0.0, cl = 0.0;
          for (int l = 0; l < n1; l++) {
            Q += R[0][i * n1 + l] * Ab[l * n2 + j];
            dQ += dRe[0][i * n1 + l] * Ab[l * n2 + j] + R[0][i * n1 + l] * dAe[l * n2 + j];
            cl -= Tc[0][i * n1 + l] * Ab[l * n2 + j];
          }
          for (int l = 0; l < n2; l++) {
            Q += Ab[i * n2 + l] * R[1][l * n2 + j];
            dQ += dAe[i * n2 + l] * R[1][l * n2 + j] + Ab[i * n2 + l] * dRe[1][l * n2 + j];
            cl += Ab[i * n2 + l] * Tc[1][l * n2 + j];
          }
          double gp, g = bdelta_ij_d(D, A, w, A->en.a[e0[0] + i], A->en.a[e0[1] + j], &gp);
          acc += cg * ((dAe[i * n2 + j] * Q + Ab[i * n2 + j] * dQ) * g + Q * cl * gp);
        }
      for (int s = 0; s < 2; s++) { free(R[s]); free(dR[s]); free(Mr[s]); free(Tc[s]); free(dRe[s]); }
      free(Ab); free(dA); free(dAe);
    }
    out[t] = acc;
  }
  return 0;
}

void orc_dm_free(orc_dm *D) {
  if (!D) return;
  for (int i = 0; i < D->nsh; i++) { arr_free(&D->acut[i]); arr_free(&D->abasis[i]); mat_free(&D->op[i]); }
  for (int i = 0; D->rho && i <= D->nshell_max; i++) mat_free(&D->rho[i]);
  for (int i = 0; D->th && i < D->nsh; i++) {
    free(D->th_off[i].a);
    for (int t = 0; t < ORC_MAX_TAN; t++) { free(D->th[i][t].a); free(D->dop[i][t].a); free(D->dz[i][t].a); }
  }
  for (int i = 0; D->drho && i <= D->nshell_max; i++)
    for (int t = 0; t < ORC_MAX_TAN; t++) free(D->drho[i][t].a);
  free(D->th); free(D->th_off); free(D->dop); free(D->dz); free(D->drho);
  free(D->acut); free(D->abasis); free(D->op); free(D->rho); free(D->trace); free(D->wn);
  arr_free(&D->site);
  free(D->omega); free(D->rhow);
  free(D);
}

int orc_dm_info(const orc_dm *D, int *nshell_max, int *spec_nshell_max, double *temp) {
  *nshell_max = D->nshell_max; *spec_nshell_max = D->spec_nshell_max; *temp = D->temp;
  return D->nsh;
}

int orc_dm_shell(const orc_dm *D, int N, orc_dm_shell_view *v) {
  if (N < 0 || N > D->nshell_max) return -1;
  const arr_t *A = &D->acut[N];
  const mat_t *R = &D->rho[N];
  v->nblocks = nblocks(A); v->nq = A->nq; v->qn = A->qn.a; v->bbe = A->bbe.a; v->en = A->en.a;
  v->nstates = A->en.n; v->rho = R->mat.a; v->rho_mbe = R->mbe.a; v->trace = D->trace[N];
  v->bbe_bc = A->bbe_bc.a; v->ev = A->ev.a; v->bbe_ev = A->bbe_ev.a;
  const mat_t *O = &D->op[N];
  v->op_nblocks = O->map.n / 2; v->op_map = O->map.a; v->op_mbe = O->mbe.a; v->op = O->mat.a;
  const arr_t *B = &D->abasis[N];
  v->basis_nstates = B->stcf.n; v->basis_stcf = B->stcf.a; v->basis_itype = B->itype.a;
  v->basis_nblocks = nblocks(B); v->basis_qn = B->qn.a; v->basis_bbe = B->bbe.a;
  v->kept = A->kept.n ? A->kept.a : NULL;
  v->wn = D->wn ? D->wn[N] : 0.0;
  return 0;
}

/* the grid is evaluated on the first call */
int orc_dm_spectrum(orc_dm *D, const double **omega, const double **rho) {
  if (!D->omega) spec_grid(D);
  *omega = D->omega; *rho = D->rhow;
  return D->nomega;
}
