// Host check of the mixed kernels' real basis (cudaprob3_amd/csrc/host_util.hpp real_basis):
// reads 18 doubles of U (row-major, re/im interleaved) and 9 of DM from stdin, prints the six
// entries of Hr and the four of row mu of W, one per line, in %.17g.
#include "../../cudaprob3_amd/csrc/host_util.hpp"

#include <cstdio>

int main() {
    double U[18], DM[9], Hr[6], Wmu[4];
    for (double& x : U)
        if (std::scanf("%lf", &x) != 1) return 2;
    for (double& x : DM)
        if (std::scanf("%lf", &x) != 1) return 2;
    cp3host::real_basis(U, DM, Hr, Wmu);
    for (double x : Hr) std::printf("%.17g\n", x);
    for (double x : Wmu) std::printf("%.17g\n", x);
    return 0;
}
