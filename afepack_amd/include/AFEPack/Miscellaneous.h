// Miscellaneous.h -- function objects and small helpers of the drop-in
// interface (include/AFEPack/Miscellaneous.h:90-233 of the reference:
// Function, FunctionFunction, innerProduct, the AFEPack namespace).
#pragma once
#include <cmath>
#include <vector>

namespace AFEPack {

template <class value_type> class Function {
public:
  virtual ~Function() = default;
  virtual value_type value(const double *) const { return value_type(0); }
  virtual std::vector<value_type> gradient(const double *) const { return {}; }
  value_type operator()(const double *p) const { return value(p); }
};

// value and gradient given by C functions (gradient optional)
template <class value_type> class FunctionFunction : public Function<value_type> {
public:
  typedef value_type (*ValuePrototype)(const double *);
  typedef std::vector<value_type> (*GradientPrototype)(const double *);
  FunctionFunction(ValuePrototype v = nullptr, GradientPrototype g = nullptr) : vf(v), gf(g) {}
  ValuePrototype valueFunction() const { return vf; }
  GradientPrototype gradientFunction() const { return gf; }
  value_type value(const double *p) const override { return vf(p); }
  std::vector<value_type> gradient(const double *p) const override { return gf(p); }

private:
  ValuePrototype vf;
  GradientPrototype gf;
};

template <class T> inline T innerProduct(const std::vector<T> &a, const std::vector<T> &b) {
  T s = 0;
  for (size_t i = 0; i < a.size(); i++) s += a[i] * b[i];
  return s;
}

template <int DIM> struct ActiveElementPairIterator {
  enum State { GREAT_THAN = -1, EQUAL = 0, LESS_THAN = 1 };
};

} // namespace AFEPack
