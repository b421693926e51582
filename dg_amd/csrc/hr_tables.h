// The three transforms of the one-launch 2D hierarchical limiter for the
// tensor Lagrange element in Gmsh node order, p = 1 and p = 2, as
// compile-time constants: Lag2MonoX (nodal values to Taylor-x coefficients
// per y slice), MonoX2MonoY (to Taylor-y per x slice) and MonoY2Lag (back to
// nodal values).  Column-major, M[r + l*Ns], the layout of the tables the
// mesh builder makes (the reference's Lag2MonoX / MonoX2MonoY / MonoY2Lag,
// handed to the Limiting constructor, limiting/limiting.h:249-290).
//
// Each line below is one column l.  Most entries are exact zeros and the
// rest are small dyadic numbers, so a kernel that knows them at compile time
// skips the zero terms and the multiplications by +-1.  For finite inputs
// that gives the same bits as the dense product y[r] = sum_l x[l]*M[r + l*Ns]
// accumulated from +0.0 in order: a +-0 product added to a non-zero sum
// leaves it unchanged, and in round-to-nearest such a sum is never -0, to
// which +0 + (-0) = +0 would apply.  An Inf or NaN input differs: the dense
// product spreads 0*Inf = NaN to every output of its column, while the
// skipped zero terms leave the outputs that do not depend on that input
// finite (a blown-up state then comes out partly finite, where the
// reference's gemm and the two-sweep path give NaN).  The host compares the built tables with these entry by entry
// (hr_tables_match) before it selects that kernel.
#pragma once

namespace dgx {

template <int P>
struct HrTables2D;

template <> struct HrTables2D<1> {
  static constexpr int Ns = 4;
  static constexpr double L2MX[16] = {
      0.5, -0.5, 0.0, 0.0,
      0.5, 0.5, 0.0, 0.0,
      0.0, 0.0, 0.5, 0.5,
      0.0, 0.0, 0.5, -0.5};
  static constexpr double MX2MY[16] = {
      0.5, -0.5, 0.5, -0.5,
      -0.5, 0.5, 0.5, -0.5,
      0.5, 0.5, 0.5, 0.5,
      -0.5, -0.5, 0.5, 0.5};
  static constexpr double MY2L[16] = {
      1.0, 0.0, 0.0, 1.0,
      -1.0, 0.0, 0.0, 1.0,
      0.0, 1.0, 1.0, 0.0,
      0.0, -1.0, 1.0, 0.0};
};

template <> struct HrTables2D<2> {
  static constexpr int Ns = 9;
  static constexpr double L2MX[81] = {
      0.0, -0.5, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0,
      0.0, 0.5, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0,
      0.0, 0.0, 0.0, 0.0, 0.5, 1.0, 0.0, 0.0, 0.0,
      0.0, 0.0, 0.0, 0.0, -0.5, 1.0, 0.0, 0.0, 0.0,
      1.0, 0.0, -2.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0,
      0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.5, 1.0,
      0.0, 0.0, 0.0, 1.0, 0.0, -2.0, 0.0, 0.0, 0.0,
      0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, -0.5, 1.0,
      0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0, -2.0};
  static constexpr double MX2MY[81] = {
      0.0, -0.5, 1.0, 0.0, -0.5, 1.0, 0.0, -0.5, 1.0,
      0.0, 0.5, -1.0, 0.0, -0.5, 1.0, 0.0, 0.0, 0.0,
      0.0, -0.25, 0.5, 0.0, -0.25, 0.5, 0.0, 0.0, 0.0,
      0.0, 0.5, 1.0, 0.0, 0.5, 1.0, 0.0, 0.5, 1.0,
      0.0, -0.5, -1.0, 0.0, 0.5, 1.0, 0.0, 0.0, 0.0,
      0.0, 0.25, 0.5, 0.0, 0.25, 0.5, 0.0, 0.0, 0.0,
      1.0, 0.0, -2.0, 1.0, 0.0, -2.0, 1.0, 0.0, -2.0,
      -1.0, 0.0, 2.0, 1.0, 0.0, -2.0, 0.0, 0.0, 0.0,
      0.5, 0.0, -1.0, 0.5, 0.0, -1.0, 0.0, 0.0, 0.0};
  static constexpr double MY2L[81] = {
      1.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0, 0.0,
      -1.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0,
      0.5, 0.0, 0.0, 0.5, 0.0, 0.0, 0.0, 0.0, 0.0,
      0.0, 1.0, 1.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0,
      0.0, -1.0, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0,
      0.0, 0.5, 0.5, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0,
      0.0, 0.0, 0.0, 0.0, 1.0, 0.0, 1.0, 0.0, 1.0,
      0.0, 0.0, 0.0, 0.0, -1.0, 0.0, 1.0, 0.0, 0.0,
      0.0, 0.0, 0.0, 0.0, 0.5, 0.0, 0.5, 0.0, 0.0};
};

// true when the three tables equal the constants above bit for bit
inline bool hr_tables_match(int P, const double* l2mx, const double* mx2my, const double* my2l) {
  auto same = [](const double* a, const double* b, int n) {
    for (int i = 0; i < n; ++i)
      if (!(a[i] == b[i]) || (a[i] == 0.0 && __builtin_signbit(a[i]) != __builtin_signbit(b[i]))) return false;
    return true;
  };
  if (P == 1) return same(l2mx, HrTables2D<1>::L2MX, 16) && same(mx2my, HrTables2D<1>::MX2MY, 16) &&
                     same(my2l, HrTables2D<1>::MY2L, 16);
  if (P == 2) return same(l2mx, HrTables2D<2>::L2MX, 81) && same(mx2my, HrTables2D<2>::MX2MY, 81) &&
                     same(my2l, HrTables2D<2>::MY2L, 81);
  return false;
}

}  // namespace dgx
