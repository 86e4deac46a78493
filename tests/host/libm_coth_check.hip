// Host-side check of the quad4 build's cosh and sinh (fdv_device.h:
// fdlibm_expm1, libm_sinh, libm_cosh), compiled with hipcc and run on the
// CPU (the functions are __host__ __device__): on random arguments spread
// logarithmically over the ranges the restatement covers, each must equal
// the host libm's expm1 / sinh / cosh bit for bit.  Prints the mismatch
// counts and exits 1 on any mismatch.
#include "../../cfdv_amd/csrc/fdv_device.h"
#include <cmath>
#include <cstdio>
#include <random>
using namespace cfdv;

int main() {
    std::mt19937_64 rng(20261204);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    long n = 0, bad_e = 0, bad_s = 0, bad_c = 0;
    for (long i = 0; i < 4000000; i++) {
        // |x| from 1e-14 to 22 (sinh, expm1) and below ln2/2 (cosh)
        const double mag = std::pow(10.0, -14.0 + 15.34 * U(rng));
        const double x = (U(rng) < 0.5) ? -mag : mag;
        n++;
        volatile double a = std::expm1(x), b = fdlibm_expm1(x);
        if (a != b && bad_e++ < 5) std::printf("expm1(%a): libm %a restated %a\n", x, a, b);
        volatile double c = std::sinh(x), d = libm_sinh(x);
        if (c != d && bad_s++ < 5) std::printf("sinh(%a): libm %a restated %a\n", x, c, d);
        const double xc = std::fmod(x, 0.3465);
        volatile double e = std::cosh(xc), f = libm_cosh(xc);
        if (e != f && bad_c++ < 5) std::printf("cosh(%a): libm %a restated %a\n", xc, e, f);
    }
    std::printf("%ld arguments: mismatches expm1 %ld, sinh %ld, cosh %ld\n", n, bad_e, bad_s, bad_c);
    return (bad_e || bad_s || bad_c) ? 1 : 0;
}
