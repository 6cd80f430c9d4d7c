// Driver for Montjoie's acoustic time-domain solver, compiled from the
// reference sources (see Makefile, target "ref").  Same call sequence as the
// reference's own mains (Program/Beginner/acous2D.cc:17-45, acous3D.cc):
// read the data file, build HyperbolicProblem<AcousticEquation<Dim>> and
// RunAll.  The Laplace matrix-free product bodies are pulled in from
// Elliptic/Helmholtz/ProdMatVectHelmholtz.cxx, as the reference's compiled
// library unit lib/Compil/Elliptic/Helmholtz/prod_var_helmholtz.cpp:12 does.
// Test infrastructure only (fixture generation, CPU baseline).
#include "MontjoieFlag.hxx"
#include "Hyperbolic/Acoustic/MontjoieAcoustic.hxx"
#include "Elliptic/Helmholtz/ProdMatVectHelmholtz.cxx"
using namespace Montjoie;
int main(int argc, char **argv)
{
  InitMontjoie(argc, argv);
  if (argc < 3) { cout << "usage: mjref_acous.x data.ini 2|3" << endl; return 1; }
  string f = string(argv[1]); string te, teq;
  getElement_Equation(f, te, teq);
  if (argv[2][0] == '2') { HyperbolicProblem<AcousticEquation<Dimension2> > V; V.RunAll(f, te, teq); }
  else { HyperbolicProblem<AcousticEquation<Dimension3> > V; V.RunAll(f, te, teq); }
  return FinalizeMontjoie();
}
