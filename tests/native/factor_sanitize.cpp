// Host driver for the sanitizer build of jja_factor.cpp (AddressSanitizer and
// UndefinedBehaviorSanitizer): factorises grid Laplacians of several sizes and
// shapes, checks L L^T = K entry by entry, and exercises the refusals (a matrix
// that is not positive definite, malformed inputs). Exit status 0 when every
// check holds; the sanitizers abort on any invalid access or undefined
// behaviour.
#include <cmath>
#include <cstdio>
#include <vector>

#include "../../include/jja.h"

namespace {

struct Csc {
  int n = 0;
  std::vector<int64_t> p;
  std::vector<int32_t> i;
  std::vector<double> x;
};

// 5-point Laplacian of an nx x ny grid plus shift * I (both triangles).
Csc grid(int nx, int ny, double shift) {
  Csc a;
  a.n = nx * ny;
  a.p.push_back(0);
  for (int c = 0; c < a.n; ++c) {
    const int cx = c % nx, cy = c / nx;
    const int nb[5] = {cy > 0 ? c - nx : -1, cx > 0 ? c - 1 : -1, c, cx + 1 < nx ? c + 1 : -1,
                       cy + 1 < ny ? c + nx : -1};
    for (int r : nb) {
      if (r < 0) continue;
      a.i.push_back(r);
      a.x.push_back(r == c ? 4.0 + shift : -1.0);
    }
    a.p.push_back(static_cast<int64_t>(a.i.size()));
  }
  return a;
}

double entry(const Csc& a, int r, int c) {
  for (int64_t q = a.p[c]; q < a.p[c + 1]; ++q)
    if (a.i[q] == r) return a.x[q];
  return 0.0;
}

int check(const Csc& a) {
  jja_factor* f = nullptr;
  int64_t nnz = 0;
  if (jja_cholesky(a.n, a.p.data(), a.i.data(), a.x.data(), &f, &nnz) != JJA_OK) {
    std::printf("factor failed: %s\n", jja_factor_last_error());
    return 1;
  }
  std::vector<int64_t> lp(a.n + 1);
  std::vector<int32_t> li(nnz);
  std::vector<double> lx(nnz);
  jja_factor_get(f, lp.data(), li.data(), lx.data());
  jja_factor_free(f);
  std::vector<double> dense(static_cast<size_t>(a.n) * a.n, 0.0);  // L, column-major
  for (int c = 0; c < a.n; ++c)
    for (int64_t q = lp[c]; q < lp[c + 1]; ++q) dense[static_cast<size_t>(c) * a.n + li[q]] = lx[q];
  double err = 0.0;
  for (int r = 0; r < a.n; ++r)
    for (int c = 0; c <= r; ++c) {
      double s = 0.0;
      for (int k = 0; k <= c; ++k)
        s += dense[static_cast<size_t>(k) * a.n + r] * dense[static_cast<size_t>(k) * a.n + c];
      err = std::fmax(err, std::fabs(s - entry(a, r, c)));
    }
  if (err > 1e-12) {
    std::printf("L L^T differs from K by %g (n=%d)\n", err, a.n);
    return 1;
  }
  return 0;
}

}  // namespace

int main() {
  int bad = 0;
  for (int nx : {1, 2, 3, 7, 16})
    for (int ny : {1, 5, 9}) bad += check(grid(nx, ny, 0.0));
  // not positive definite: shifted far below zero
  Csc neg = grid(4, 4, -10.0);
  jja_factor* f = nullptr;
  int64_t nnz = 0;
  if (jja_cholesky(neg.n, neg.p.data(), neg.i.data(), neg.x.data(), &f, &nnz) != JJA_EFACTOR) {
    std::printf("an indefinite matrix was not refused\n");
    ++bad;
  }
  // malformed: negative order, row index out of range
  if (jja_cholesky(-1, neg.p.data(), neg.i.data(), neg.x.data(), &f, &nnz) == JJA_OK) ++bad;
  Csc broken = grid(3, 3, 0.0);
  broken.i[2] = 99;
  if (jja_cholesky(broken.n, broken.p.data(), broken.i.data(), broken.x.data(), &f, &nnz) == JJA_OK) {
    std::printf("an out-of-range row index was not refused\n");
    ++bad;
  }
  std::printf(bad ? "FAILED\n" : "sanitized factor checks passed\n");
  return bad ? 1 : 0;
}
