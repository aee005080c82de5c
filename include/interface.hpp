// interface.hpp -- header-only drop-in for defmm's include/interface.hpp
// (reference include/interface.hpp:45-224): the same defmm::IBFMM_Mat<3>
// surface (add*Particle*, prcmpt, gemv, CheckFULL) and the global names the
// reference's test uses (flt, cplx, Vecc, cplx0, urand, f_cyan, f_def), backed
// by the MI355X evaluation pass through the C ABI of defmm_amd.h.
//
// Differences: every function is inline (safe in several translation units);
// a plan may be evaluated any number of times; errors are reported on stderr
// and exit(1) as the reference does, with the library's message.
// Link with -ldefmm_amd (defmm_amd/libdefmm_amd.so).
#ifndef DEFMM_AMD_INTERFACE_HPP
#define DEFMM_AMD_INTERFACE_HPP

#include <cmath>
#include <complex>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <vector>

#include "defmm_amd.h"

typedef double flt;
typedef std::complex<flt> cplx;
const cplx cplx0 = cplx(0., 0.);
const flt pi4 = 4. * M_PI;
#ifndef urand
#define urand rand() / flt(RAND_MAX)
#endif

namespace gaia {
// Minimal vector of the reference (gaia_vectors.hpp:16-173): owning,
// fixed-size, indexable.
template <typename T>
class Vec {
  std::vector<T> d_;

 public:
  Vec() {}
  explicit Vec(int n) : d_(n) {}
  Vec(int n, const T& v) : d_(n, v) {}
  const Vec& operator=(const T& v) {
    for (auto& x : d_) x = v;
    return *this;
  }
  T& operator[](int i) { return d_[i]; }
  const T& operator[](int i) const { return d_[i]; }
  friend T* K_(Vec& v) { return v.d_.data(); }
  friend int Size(const Vec& v) { return (int)v.d_.size(); }
};
}  // namespace gaia
typedef gaia::Vec<cplx> Vecc;

// ANSI colour manipulators used by the reference's test (gaia_colors.hpp).
inline std::ostream& f_cyan(std::ostream& s) { return s << "\033[36m"; }
inline std::ostream& f_def(std::ostream& s) { return s << "\033[39m"; }

namespace defmm {

inline void defmm_check(int rc, const char* what) {
  if (rc != DEFMM_OK) {
    std::cerr << "defmm: " << what << ": " << defmm_strerror(rc) << ": " << defmm_last_error()
              << std::endl;
    std::exit(1);
  }
}

template <int DIM>
class IBFMM_Mat;

template <>
class IBFMM_Mat<3> {
  std::vector<flt> src_, trg_;
  defmm_plan plan_ = nullptr;
  int precision_ = DEFMM_FP64;

 public:
  IBFMM_Mat() {}
  explicit IBFMM_Mat(int precision) : precision_(precision) {}
  ~IBFMM_Mat() {
    if (plan_) defmm_plan_destroy(plan_);
  }
  IBFMM_Mat(const IBFMM_Mat&) = delete;
  IBFMM_Mat& operator=(const IBFMM_Mat&) = delete;

  void addSourceParticle(flt part[3]) { src_.insert(src_.end(), part, part + 3); }
  void addTargetParticle(flt part[3]) { trg_.insert(trg_.end(), part, part + 3); }
  void addSourceParticlesINP(const char* name, int nparts) { readINP(name, nparts, src_); }
  void addTargetParticlesINP(const char* name, int nparts) { readINP(name, nparts, trg_); }

  void prcmpt(int L, int Ncrit, flt kappa, int lvMAX = -1) {
    if (plan_) defmm_plan_destroy(plan_);
    plan_ = nullptr;
    defmm_check(defmm_plan_create(src_.data(), (int64_t)(src_.size() / 3), trg_.data(),
                                  (int64_t)(trg_.size() / 3), L, Ncrit, kappa, lvMAX,
                                  precision_, &plan_),
                "prcmpt");
    kappa_ = kappa;
  }

  friend void gemv(IBFMM_Mat<3>& A, Vecc& In, Vecc& Out) {
    if (!A.plan_) defmm_check(DEFMM_ESTATE, "gemv before prcmpt");
    A.in_.assign((const flt*)K_(In), (const flt*)K_(In) + 2 * Size(In));
    defmm_check(defmm_execute(A.plan_, (const double*)K_(In), (double*)K_(Out)), "gemv");
    A.out_.assign((const flt*)K_(Out), (const flt*)K_(Out) + 2 * Size(Out));
  }

  // FP64 direct sum at every target (on the GPU) against the last gemv:
  // relative L-inf, L1, L2 errors (interface.hpp:178-221).
  friend void CheckFULL(IBFMM_Mat<3>& A) {
    int nt = (int)(A.trg_.size() / 3);
    std::vector<cplx> ex(nt);
    defmm_check(defmm_direct(A.src_.data(), (int64_t)(A.src_.size() / 3), A.in_.data(),
                             A.trg_.data(), nt, A.kappa_, (double*)ex.data()),
                "CheckFULL");
    double a = 0, b = 0, c1 = 0, d1 = 0, c2 = 0, d2 = 0;
    for (int i = 0; i < nt; i++) {
      cplx p(A.out_[2 * i], A.out_[2 * i + 1]);
      double e = std::abs(ex[i] - p), m = std::abs(ex[i]);
      a = std::max(a, e);
      b = std::max(b, m);
      c1 += e;
      d1 += m;
      c2 += e * e;
      d2 += m * m;
    }
    std::cout << "Relative L^oo norm: " << f_cyan << a / b << f_def << std::endl;
    std::cout << "Relative L^1  norm: " << f_cyan << c1 / d1 << f_def << std::endl;
    std::cout << "Relative L^2  norm: " << f_cyan << std::sqrt(c2 / d2) << f_def << std::endl;
  }

 private:
  flt kappa_ = 0;
  std::vector<flt> in_, out_;
  static void readINP(const char* name, int n, std::vector<flt>& dst) {
    std::ifstream f(name);
    dst.assign(3 * (size_t)n, 0.);
    for (size_t i = 0; i < 3 * (size_t)n; i++) f >> dst[i];
  }
};

}  // namespace defmm

#endif
