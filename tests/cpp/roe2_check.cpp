// roe2_check.cpp: roe_local2 (the 2-D solver's Roe flux with one tangential
// velocity) against roe_local (the 3-D routine) with the third velocity
// component zero, in either tangential slot, on random face states in FP64
// and FP32.  Prints the number of states whose outputs differ in value.
#include <cstdio>
#include <random>
#include "ox_physics.hpp"
using namespace ox;
int main() {
  ox_config c3{}; c3.nx=c3.ny=c3.nz=8; c3.order=2; c3.Lx=c3.Ly=c3.Lz=1; c3.gamma=1.4; c3.R=287; c3.Pr=0.72; c3.Prt=0.9; c3.threshold=0.4;
  HostConsts hc; hc.init(c3);
  PhysC<double> c = make_physc<double>(hc);
  PhysC<float> cf = make_physc<float>(hc);
  std::mt19937_64 g(1); std::uniform_real_distribution<double> U(-1,1);
  long bad=0, badf=0, n=0;
  for (int it=0; it<300000; ++it) {
    double r1=1+0.5*U(g), r2=1+0.5*U(g), u1=U(g)*3, u2=U(g)*3, t1=U(g)*3, t2=U(g)*3, p1=1+0.5*U(g), p2=1+0.5*U(g);
    if (it%7==0) { t2=t1; } if (it%11==0) { u2=u1; r2=r1; } if (it%13==0) {t1=0; t2=0;}
    double o3[5], o2[4], q3[5];
    roe_local<double>(c, r1,r2,u1,u2,t1,t2,0.0,0.0,p1,p2,o3);
    roe_local2<double>(c, r1,r2,u1,u2,t1,t2,p1,p2,o2);
    roe_local<double>(c, r1,r2,u1,u2,0.0,0.0,t1,t2,p1,p2,q3);
    bool ok = o3[0]==o2[0] && o3[1]==o2[1] && o3[2]==o2[2] && o3[4]==o2[3] && q3[0]==o2[0] && q3[1]==o2[1] && q3[3]==o2[2] && q3[4]==o2[3];
    if (!ok) { if (bad<5) printf("bad %d: %.17g %.17g | %.17g %.17g | %.17g %.17g %.17g\n", it, o3[2], o2[2], o3[4], o2[3], q3[3], q3[4], q3[0]-o2[0]); ++bad; }
    float f3[5], f2[4], g3[5];
    roe_local<float>(cf, (float)r1,(float)r2,(float)u1,(float)u2,(float)t1,(float)t2,0.f,0.f,(float)p1,(float)p2,f3);
    roe_local2<float>(cf, (float)r1,(float)r2,(float)u1,(float)u2,(float)t1,(float)t2,(float)p1,(float)p2,f2);
    roe_local<float>(cf, (float)r1,(float)r2,(float)u1,(float)u2,0.f,0.f,(float)t1,(float)t2,(float)p1,(float)p2,g3);
    bool okf = f3[0]==f2[0] && f3[1]==f2[1] && f3[2]==f2[2] && f3[4]==f2[3] && g3[3]==f2[2] && g3[1]==f2[1];
    if (!okf) ++badf;
    ++n;
  }
  printf("n=%ld bad=%ld badf=%ld\n", n, bad, badf);
  return (bad || badf) ? 1 : 0;
}
