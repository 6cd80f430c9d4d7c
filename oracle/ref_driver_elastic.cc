// Driver for Montjoie's isotropic elastodynamics solver, compiled from the
// reference sources (see Makefile, target "ref").  It is the same call
// sequence as the reference's own mains (Program/Beginner/elastic3D.cc:11-40,
// elastic2D.cc): read the data file, build
// HyperbolicProblem<TimeElasticEquation<Dim>> and RunAll; the dimension is
// the optional second argument (3 by default).  Test infrastructure only
// (fixture generation, CPU baseline).
#include "MontjoieFlag.hxx"
#include "Hyperbolic/Elastic/MontjoieTimeElastic.hxx"
using namespace Montjoie;
int main(int argc, char **argv)
{
  InitMontjoie(argc, argv);
  if (argc < 2) { cout << "usage: mjref_elastic.x data.ini [2|3]" << endl; return 1; }
  string f = string(argv[1]); string te, teq;
  getElement_Equation(f, te, teq);
  if (argc > 2 && argv[2][0] == '2') {
    HyperbolicProblem<TimeElasticEquation<Dimension2> > V;
    V.RunAll(f, te, teq);
  } else {
    HyperbolicProblem<TimeElasticEquation<Dimension3> > V;
    V.RunAll(f, te, teq);
  }
  return FinalizeMontjoie();
}
