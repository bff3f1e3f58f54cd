// Host-side sparse Cholesky factorisation of the face system matrix K.
//
// The reference factorises K = A (L + diag(1/(Rv+Cv))) A^T once per run with
// scipy's SuperLU (time_evolution.py:504-506). K is symmetric positive definite,
// so this library computes K = L L^T instead, in the row order chosen by the
// caller (nested dissection), with an up-looking algorithm:
//
//   * elimination tree by ancestor path compression;
//   * for row k, the nonzero pattern of L(k, :) is the set of tree nodes reached
//     walking up from every i < k with K(i, k) != 0 (the "row subtree");
//   * row k is then a sparse triangular solve against the columns computed so
//     far, followed by the diagonal sqrt.
//
// A first pass over the row subtrees gives exact column counts so the numeric
// pass writes L in compressed-column form without reallocation. Columns are
// produced with ascending row indices, diagonal first.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/jja.h"

namespace {

thread_local std::string g_factor_error;

struct Factor {
  int32_t n = 0;
  std::vector<int64_t> colptr;
  std::vector<int32_t> rowidx;
  std::vector<double> val;
};

// Walk the row subtree of row k: push nodes reached from the upper entries of
// column k into stack[top..n) in topological order (descendants first).
int32_t row_subtree(int32_t k, const int64_t* Ap, const int32_t* Ai,
                    const int32_t* parent, int32_t* stack, int32_t* mark, int32_t stamp,
                    int32_t n) {
  int32_t top = n;
  mark[k] = stamp;
  for (int64_t p = Ap[k]; p < Ap[k + 1]; ++p) {
    int32_t i = Ai[p];
    if (i > k) continue;
    int32_t len = 0;
    // climb until a marked node; collect the path in stack[0..len)
    while (mark[i] != stamp) {
      stack[len++] = i;
      mark[i] = stamp;
      i = parent[i];
      if (i < 0) break;
    }
    while (len > 0) stack[--top] = stack[--len];
  }
  return top;
}

}  // namespace

extern "C" {

const char* jja_factor_last_error(void) { return g_factor_error.c_str(); }

int jja_cholesky(int32_t n, const int64_t* Ap, const int32_t* Ai, const double* Ax,
                 jja_factor** out, int64_t* nnz_out) {
  if (!out || !nnz_out || n < 0 || (n > 0 && (!Ap || !Ai || !Ax))) {
    g_factor_error = "jja_cholesky: invalid arguments";
    return JJA_EINVAL;
  }
  Factor* F = new (std::nothrow) Factor();
  if (!F) {
    g_factor_error = "jja_cholesky: out of host memory";
    return JJA_ENOMEM;
  }
  F->n = n;
  try {
    std::vector<int32_t> parent(n, -1), ancestor(n, -1);
    // elimination tree (upper part of each column)
    for (int32_t k = 0; k < n; ++k) {
      for (int64_t p = Ap[k]; p < Ap[k + 1]; ++p) {
        int32_t i = Ai[p];
        if (i < 0 || i >= n) {
          g_factor_error = "jja_cholesky: row index out of range";
          delete F;
          return JJA_EINVAL;
        }
        while (i != -1 && i < k) {
          int32_t next = ancestor[i];
          ancestor[i] = k;
          if (next == -1) {
            parent[i] = k;
            break;
          }
          i = next;
        }
      }
    }
    // column counts from row subtrees
    std::vector<int32_t> stack(n), mark(n, -1);
    std::vector<int64_t> count(n, 1);
    for (int32_t k = 0; k < n; ++k) {
      int32_t top = row_subtree(k, Ap, Ai, parent.data(), stack.data(), mark.data(), k, n);
      for (int32_t t = top; t < n; ++t) count[stack[t]]++;
    }
    F->colptr.assign(n + 1, 0);
    for (int32_t j = 0; j < n; ++j) F->colptr[j + 1] = F->colptr[j] + count[j];
    const int64_t nnz = F->colptr[n];
    F->rowidx.assign(nnz, 0);
    F->val.assign(nnz, 0.0);
    std::vector<int64_t> next(F->colptr.begin(), F->colptr.end() - 1);
    std::vector<double> x(n, 0.0);
    std::fill(mark.begin(), mark.end(), -1);
    for (int32_t k = 0; k < n; ++k) {
      int32_t top = row_subtree(k, Ap, Ai, parent.data(), stack.data(), mark.data(), k, n);
      double d = 0.0;
      for (int64_t p = Ap[k]; p < Ap[k + 1]; ++p) {
        int32_t i = Ai[p];
        if (i < k) x[i] += Ax[p];
        else if (i == k) d += Ax[p];
      }
      for (int32_t t = top; t < n; ++t) {
        const int32_t j = stack[t];
        const double lkj = x[j] / F->val[F->colptr[j]];
        x[j] = 0.0;
        for (int64_t p = F->colptr[j] + 1; p < next[j]; ++p) x[F->rowidx[p]] -= F->val[p] * lkj;
        d -= lkj * lkj;
        F->rowidx[next[j]] = k;
        F->val[next[j]] = lkj;
        next[j]++;
      }
      if (!(d > 0.0) || !std::isfinite(d)) {
        g_factor_error = "jja_cholesky: matrix is not positive definite (pivot " +
                         std::to_string(k) + ")";
        delete F;
        return JJA_EFACTOR;
      }
      F->rowidx[next[k]] = k;
      F->val[next[k]] = std::sqrt(d);
      next[k]++;
    }
  } catch (const std::bad_alloc&) {
    g_factor_error = "jja_cholesky: out of host memory";
    delete F;
    return JJA_ENOMEM;
  }
  *out = reinterpret_cast<jja_factor*>(F);
  *nnz_out = F->colptr[n];
  return JJA_OK;
}

int jja_factor_get(const jja_factor* f, int64_t* colptr, int32_t* rowidx, double* val) {
  const Factor* F = reinterpret_cast<const Factor*>(f);
  if (!F || !colptr || !rowidx || !val) {
    g_factor_error = "jja_factor_get: invalid arguments";
    return JJA_EINVAL;
  }
  std::memcpy(colptr, F->colptr.data(), sizeof(int64_t) * (F->n + 1));
  std::memcpy(rowidx, F->rowidx.data(), sizeof(int32_t) * F->rowidx.size());
  std::memcpy(val, F->val.data(), sizeof(double) * F->val.size());
  return JJA_OK;
}

void jja_factor_free(jja_factor* f) { delete reinterpret_cast<Factor*>(f); }

}  // extern "C"
