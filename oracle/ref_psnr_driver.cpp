// Main program around the reference's own RNG sources (src/PSNR/PSNR.cpp and
// src/PSNR/ran3.c), compiled where they lie by oracle/Makefile into
// oracle/_ref/ref_psnr.  Test infrastructure only: it prints, as JSON, the
// state the reference host hands to the kernel (PSNR::InitRand3,
// GetRan3Array, GetInextInextp; CalculatePropagation.cpp:1893-1899) and the
// integer streams of PSNR::GetRan3 and of Numerical Recipes ran3(), for the
// seeds named on the command line.
#include "PSNR.hpp"
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

extern "C" double ran3(long *idum);

int main(int argc, char **argv)
{
    int ndraw = 200;
    std::printf("[\n");
    for (int a = 1; a < argc; ++a) {
        int seed = std::atoi(argv[a]);
        PSNR g;
        g.SetSeed(seed);
        g.InitRand3();
        std::vector<float> arr;
        std::vector<int> idx;
        g.GetRan3Array(arr);
        g.GetInextInextp(idx);
        std::printf("{\"seed\": %d, \"inext\": %d, \"inextp\": %d,\n \"ran3_array\": [", seed, idx[0], idx[1]);
        for (int i = 1; i < 56; ++i) std::printf("%s%.1f", i > 1 ? ", " : "", (double)arr[i]);
        std::printf("],\n \"psnr_draws\": [");
        for (int i = 0; i < ndraw; ++i) {
            double r = g.GetRan3();
            std::printf("%s%ld", i ? ", " : "", (long)std::llround(r * 1e9));
        }
        // ran3.c keeps static state: a negative idum re-initialises it.
        long idum = -std::labs((long)seed);
        std::printf("],\n \"ran3c_draws\": [");
        for (int i = 0; i <= ndraw; ++i) {
            double r = ran3(&idum);
            std::printf("%s%ld", i ? ", " : "", (long)std::llround(r * 1e9));
        }
        std::printf("]}%s\n", a + 1 < argc ? "," : "");
    }
    std::printf("]\n");
    return 0;
}
