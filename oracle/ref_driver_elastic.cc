// Driver for Montjoie's 3-D isotropic elastodynamics solver, compiled from the
// reference sources (see Makefile, target "ref").  It is the same call
// sequence as the reference's own main (Program/Beginner/elastic3D.cc:11-40):
// read the data file, build HyperbolicProblem<TimeElasticEquation<Dimension3>>
// and RunAll.  Test infrastructure only (fixture generation).
#include "MontjoieFlag.hxx"
#include "Hyperbolic/Elastic/MontjoieTimeElastic.hxx"
using namespace Montjoie;
int main(int argc, char **argv)
{
  InitMontjoie(argc, argv);
  if (argc < 2) { cout << "usage: mjref_elastic.x data.ini" << endl; return 1; }
  string f = string(argv[1]); string te, teq;
  getElement_Equation(f, te, teq);
  HyperbolicProblem<TimeElasticEquation<Dimension3> > V;
  V.RunAll(f, te, teq);
  return FinalizeMontjoie();
}
