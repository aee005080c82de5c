// The reference's test program (test/testinterface.cpp of the reference)
// written against the drop-in header: refined-cube mesh as sources and
// targets, rand()/RAND_MAX real charges, prcmpt(ORDER, NCRIT, 1), gemv,
// CheckFULL.  argv: mesh path, ORDER, NCRIT.
#include "interface.hpp"

using namespace defmm;

int main(int argc, char* argv[]) {
  const char* mesh = argc > 1 ? argv[1] : "./mesh/refinedcube.inp";
  const int ORDER = argc > 2 ? atoi(argv[2]) : 6;
  const int NCRIT = argc > 3 ? atoi(argv[3]) : 32;
  const flt KAPPA = 1.;
  IBFMM_Mat<3> A;
  int N = 41334;
  A.addSourceParticlesINP(mesh, N);
  A.addTargetParticlesINP(mesh, N);
  std::cout << "Refined cube test case with N=" << f_cyan << N << f_def << std::endl;
  Vecc In(N), Out(N);
  Out = cplx0;
  for (int n = 0; n < N; n++) In[n] = cplx(urand);
  A.prcmpt(ORDER, NCRIT, KAPPA);
  gemv(A, In, Out);
  CheckFULL(A);
  return 0;
}
